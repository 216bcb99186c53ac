"""Writes the known-answer fixtures of tests/golden/.

Every expected value here is worked by hand from the reference's source
semantics (not produced by running any implementation):

* METCF, ROW_WINDOW 8 and 16, of a 10x10 matrix
  (include/format.hpp:215-274: sorted unique columns per window padded with
  UINT32_MAX, bit (r % RW)*8 + c % 8, values grouped per tile in CSR order,
  tcOffset an exclusive prefix with a leading 0)
* AdpBME group boundaries (include/format.hpp:346,366-382: first tile of a
  window costs 3+1+2, others 3+1, target 16*(3+1)+2 = 66)
* balanceStrategy IBD (include/utils.hpp:75-85: windows 0..W-2 only)
* Matrix Market expansion order (include/mmio.hpp:133-158) and banner
  failures (:37-69)

Run: python tests/golden/make_golden.py
"""
import json
from pathlib import Path

HERE = Path(__file__).resolve().parent
M = 4294967295

# 10x10 matrix, canonical rows (col: value)
A10 = {0: {0: 1, 3: 2, 9: 3}, 1: {1: 4}, 2: {3: 5, 5: 6}, 3: {},
       4: {0: 7, 2: 8, 4: 9, 6: 10, 8: 11}, 5: {7: 12}, 6: {9: 13}, 7: {1: 14},
       8: {8: 15, 9: 16}, 9: {0: 17}}


def csr_of(rows: dict, n: int):
    rp, ci, va = [0], [], []
    for r in range(n):
        for c in sorted(rows.get(r, {})):
            ci.append(c)
            va.append(rows[r][c])
        rp.append(len(ci))
    return rp, ci, va


def main():
    rp, ci, va = csr_of(A10, 10)
    kat = {
        "a10": {"rows": 10, "cols": 10, "row_ptr": rp, "col_idx": ci, "val": va},
        # window 0 (rows 0-7): columns 0..9 -> 2 tiles; window 1 (rows 8-9):
        # columns {0,8,9} -> 1 tile.  Tile 0 bits {0,3,9,19,21,32,34,36,38,47,57},
        # tile 1 bits {1,32,49}, tile 2 bits {1,2,8}.
        "metcf_rw8": {
            "rowWindowOffset": [0, 2, 3],
            "sparseA2B": [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, M, M, M, M, M, M,
                          0, 8, 9, M, M, M, M, M],
            "tcOffset": [0, 11, 14, 17],
            "tcLocalBit": [0x0200805500280209, 0x0002000100000002, 0x0000000000000106],
            "data": [1, 2, 4, 5, 6, 7, 8, 9, 10, 12, 14, 3, 11, 13, 15, 16, 17],
            "groupOffset": [0, 3],
            "rowIndices": [0, 0, 8],
            "ibd": 0, "balanced": False,
        },
        # one window (rows 0-9), columns 0..9 -> 2 tiles of 128 bits.
        # Tile 0: word0 as above, word1 bit 8 (r9,c0 -> 72); 12 values.
        # Tile 1 (cond 8,9): bits {1 (r0c9), 32 (r4c8), 49 (r6c9), 64, 65}.
        "metcf_rw16": {
            "rowWindowOffset": [0, 2],
            "sparseA2B": [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, M, M, M, M, M, M],
            "tcOffset": [0, 12, 17],
            "tcLocalBit": [0x0200805500280209, 0x0000000000000100,
                           0x0002000100000002, 0x0000000000000003],
            "data": [1, 2, 4, 5, 6, 7, 8, 9, 10, 12, 14, 17, 3, 11, 13, 15, 16],
            "groupOffset": [0, 2],
            "rowIndices": [0, 0],
            "ibd": 0, "balanced": False,
        },
        # AdpBME grouping cases: (rows, per-window column count) -> groupOffset
        "adpbme_groups": [
            # one 8-row window with 160 distinct columns = 20 tiles:
            # 6 + 15*4 = 66 fits, the 17th tile would exceed -> [0,16,20]
            {"windows": [160], "groupOffset": [0, 16, 20]},
            # two windows of 10 tiles: 42 + 6 + 4*4 = 64, next 68 > 66
            {"windows": [80, 80], "groupOffset": [0, 15, 20]},
            # empty window in the middle contributes nothing
            {"windows": [8, 0, 8], "groupOffset": [0, 2]},
        ],
        # IBD: per-window tile counts -> (ibd, balanced)
        "ibd_cases": [
            {"tiles": [40, 1, 1, 1], "ibd": 12, "balanced": True},
            {"tiles": [1, 1, 1, 40], "ibd": 6, "balanced": False},
            {"tiles": [2, 1], "ibd": 0, "balanced": False},
        ],
        "mtx": {
            "symmetric_3x3": {
                "text": "%%MatrixMarket matrix coordinate real symmetric\n"
                        "% a comment\n3 3 3\n1 1 1.5\n2 1 2.5\n3 2 3.5\n",
                "coo": [[0, 0, 1.5], [1, 0, 2.5], [0, 1, 2.5], [2, 1, 3.5], [1, 2, 3.5]],
                "row_ptr": [0, 2, 4, 5], "col_idx": [0, 1, 0, 2, 1],
                "val": [1.5, 2.5, 2.5, 3.5, 3.5],
            },
            "general_pattern": {
                "text": "%%MatrixMarket matrix coordinate pattern general\n"
                        "2 3 3\n2 3\n1 2\n2 1\n",
                "coo": [[1, 2, 1.0], [0, 1, 1.0], [1, 0, 1.0]],
                "row_ptr": [0, 1, 3], "col_idx": [1, 2, 0], "val": [1.0, 1.0, 1.0],
            },
            "bad_banners": [
                "%%MatrixMarket matrix coordinate real\n1 1 1\n1 1 1\n",
                "%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 1\n",
                "%%MatrixMarket Matrix coordinate real general\n1 1 1\n1 1 1\n",
                "%%MatrixMarket matrix coordinate Real general\n1 1 1\n1 1 1\n",
                "%%MatrixMarket matrix coordinate real General\n1 1 1\n1 1 1\n",
            ],
        },
    }
    (HERE / "kat.json").write_text(json.dumps(kat, indent=1) + "\n")
    print(f"wrote {HERE / 'kat.json'}")


if __name__ == "__main__":
    main()
