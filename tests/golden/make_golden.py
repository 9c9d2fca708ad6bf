"""Writes tests/golden/kat.json: the known-answer vectors of hucsmn/psacak.

* The three ``mississippi`` doc-test vectors are literal in the crate
  (lib.rs:15, lib.rs:27, lib.rs:42) and are copied here as data.
* The eleven table texts of ``tablecheck_sacak8`` / ``tablecheck_sacak32``
  (psacak8.rs:508-524, psacak32.rs:826-842) are data too; the crate checks
  them against its brute-force ``saca_tiny`` (common.rs:13-25), so their
  expected SAs are computed here by the same brute-force definition (Python
  slice comparison, shorter prefix first).

Run: python tests/golden/make_golden.py
"""
import json
import os

TABLE = [
    [0, 0, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 1],
    [5, 4, 3, 2, 1, 0],
    [3, 4, 5, 2, 0, 1],
    [2, 0, 2, 0, 2, 1, 4, 3],
    [3, 2, 1, 3, 2, 3, 2, 1, 0, 1],
    [2, 1, 4, 1, 1, 4, 1, 3, 1],
    [2, 1, 1, 3, 3, 1, 1, 3, 3, 1, 2, 1],
    [2, 2, 1, 4, 4, 1, 4, 4, 1, 3, 3, 1, 1],
    [6, 8, 9, 5, 2, 4, 3, 0, 0, 7, 1, 2],
    [1, 2, 2, 1, 1, 0, 0, 1, 1, 2, 2, 0, 0, 2, 2, 0, 1, 0, 2, 0, 1, 1, 1, 1, 2, 2, 0, 0, 2, 1,
     2, 1, 1, 0, 2, 1, 2, 2, 0, 2, 1, 1, 2, 2, 2, 1, 2, 0, 0, 1, 2, 0, 0, 0, 1, 2, 2, 2, 1, 1,
     1, 1, 2, 0, 2, 1, 1, 1, 2, 1, 0, 1],
]


def brute_sa(t):
    return sorted(range(len(t)), key=lambda i: list(t[i:]))


def main():
    kat = {
        "doc_tests": {
            "text": "mississippi",
            "psacak": [10, 7, 4, 1, 0, 9, 8, 6, 3, 5, 2],
            "psacak_inplace": [10, 7, 4, 1, 0, 9, 8, 6, 3, 5, 2],
            "with_sentinel": [11, 10, 7, 4, 1, 0, 9, 8, 6, 3, 5, 2],
        },
        "table": [{"text": t, "sa": brute_sa(t)} for t in TABLE],
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat.json")
    with open(path, "w") as f:
        json.dump(kat, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
