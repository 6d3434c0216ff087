"""Write tests/golden/golden.json from the reference's own test data and the
definitions those tests rest on — never from the oracle.

* sort sets: the five code sets of src/bvh/sort.rs:156-175 (the Knuth set is
  integer-exact: (i * 2654435761) mod 2^32 mod 97), with the order that
  src/bvh/build.rs:87-92 defines — indices sorted by (code, index) — computed
  by Python's sort;
* Morton asserts of src/bvh/node.rs:119-125;
* binary radix trees: for small code sets (duplicates included), the tree
  Karras' construction must produce, found top-down by brute force on the
  unique 64-bit keys (code << 32 | index): each range splits after the last
  key whose highest differing bit (between the range's ends) is 0, and the
  node numbering is Karras 2012 §4 (an inner node is named by the end of its
  range that its parent split at; leaves are n - 1 + slot).

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import pathlib

OUT = pathlib.Path(__file__).resolve().parent / "golden.json"


def sort_sets():
    sets = [
        [],
        [7],
        [5] * 8,
        [9, 1, 8, 1, 7, 1, 0, 0xFFFFFFFF, 0, 3],
        [((i * 2654435761) & 0xFFFFFFFF) % 97 for i in range(1000)],
    ]
    out = []
    for codes in sets:
        order = sorted(range(len(codes)), key=lambda i: (codes[i], i))
        out.append({"codes": codes, "order": order, "sorted": [codes[i] for i in order]})
    return out


def radix_tree(codes):
    """Links [left, right, first, last] per inner node, by brute force."""
    n = len(codes)
    order = sorted(range(n), key=lambda i: (codes[i], i))
    srt = [codes[i] for i in order]
    keys = [(srt[s] << 32) | s for s in range(n)]
    links = [None] * (n - 1)

    def split(f, l):
        top = (keys[f] ^ keys[l]).bit_length() - 1
        g = f
        while g + 1 <= l and not ((keys[g + 1] >> top) & 1):
            g += 1
        return g

    stack = [(0, 0, n - 1)] if n > 1 else []
    while stack:
        node, f, l = stack.pop()
        g = split(f, l)
        left = n - 1 + g if g == f else g
        right = n + g if g + 1 == l else g + 1
        links[node] = [left, right, f, l]
        if g != f:
            stack.append((g, f, g))
        if g + 1 != l:
            stack.append((g + 1, g + 1, l))
    return {"codes": codes, "sorted": srt, "links": links}


def trees():
    sets = [
        [3, 1],
        [4, 4, 4],
        [0, 1, 2, 3, 4, 5, 6, 7],
        [7, 7, 1, 1, 3, 3, 0, 0, 5],
        [1023, 0, 511, 512, 511, 511, 2, 1],
        [((i * 2654435761) & 0xFFFFFFFF) % 13 for i in range(40)],
        [(i * 7919) % 1000 for i in range(64)],
        [0x3FFFFFFF, 0, 0x3FFFFFFF, 0x15555555, 0x2AAAAAAA, 0x15555555],
    ]
    return [radix_tree(c) for c in sets]


def main():
    data = {
        "sort_sets": sort_sets(),
        "morton": {
            "ones_code": (1 << 30) - 1,
            "order_pairs": [[[0.1, 0.1, 0.1], [0.9, 0.9, 0.9]]],
            "group_pair": [[1.0, 1.0, 1.0, 0], [0.0, 0.0, 0.0, 1]],
        },
        "radix_trees": trees(),
    }
    OUT.write_text(json.dumps(data))
    print(OUT)


if __name__ == "__main__":
    main()
