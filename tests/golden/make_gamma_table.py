"""Write gamma_table.json: the reference's sparse gamma tables as data.

The reference evaluates four sparse views of the DeGrand-Rossi matrices with
switch tables: for Gamma index g and row r it returns the column of the one
nonzero entry (gamma_index, include/gamma.cuh:58-80) and its value for

  * ``G``      gamma_data<false>          (include/gamma.cuh:82-104)
  * ``GT``     gamma_data<true>           (include/gamma.cuh:105-125)
  * ``Gg5``    gamma_gamma5_data<false>   (include/gamma.cuh:225-248)
  * ``g5GT``   gamma_gamma5_data<true>    (include/gamma.cuh:249-269)

Each table below lists, per g, the row-sign function and the constant factor
read off those switch statements; row_01/row_02/row_03 are the sign helpers of
include/gamma.cuh:41-56.  The script evaluates them into a JSON fixture of
(column, re, im) triples that tests/test_oracle_gamma.py checks the oracle's
closed-form matrices against.
"""

import json
import os

I = 1j


def one(r):
    return 1


def row_01(r):  # +1 for rows 0,1; -1 for rows 2,3
    return 1 - 2 * (r >> 1)


def row_02(r):  # +1 for rows 0,2; -1 for rows 1,3
    return 1 - 2 * (r & 1)


def row_03(r):  # +1 for rows 0,3; -1 for rows 1,2
    return 1 - 2 * ((r >> 1) ^ (r & 1))


def column(g, r):
    if g in (0, 3, 12, 15):
        return r
    if g in (1, 2, 13, 14):
        return 3 - r
    if g in (4, 7, 8, 11):
        return (r + 2) % 4
    return (5 - r) % 4


TABLES = {
    "G": [(one, 1), (row_01, I), (row_03, -1), (row_02, -I), (row_03, I), (row_02, -1), (one, -I), (row_01, 1),
          (one, 1), (row_01, I), (row_03, -1), (row_02, -I), (row_03, I), (row_02, -1), (one, -I), (row_01, 1)],
    "GT": [(one, 1), (row_01, -I), (row_03, -1), (row_02, -I), (row_03, -I), (row_02, 1), (one, -I), (row_01, -1),
           (one, 1), (row_01, I), (row_03, 1), (row_02, -I), (row_03, I), (row_02, 1), (one, -I), (row_01, 1)],
    "Gg5": [(row_01, 1), (one, -I), (row_02, 1), (row_03, -I), (row_02, -I), (row_03, -1), (row_01, -I), (one, -1),
            (row_01, -1), (one, I), (row_02, -1), (row_03, I), (row_02, I), (row_03, 1), (row_01, I), (one, 1)],
    "g5GT": [(row_01, 1), (one, I), (row_02, 1), (row_03, -I), (row_02, I), (row_03, 1), (row_01, -I), (one, 1),
             (row_01, -1), (one, I), (row_02, 1), (row_03, I), (row_02, I), (row_03, -1), (row_01, I), (one, 1)],
}


def main():
    out = {}
    for name, rows in TABLES.items():
        tab = []
        for g, (fn, coef) in enumerate(rows):
            entries = []
            for r in range(4):
                v = complex(fn(r) * coef)
                entries.append([column(g, r), v.real, v.imag])
            tab.append(entries)
        out[name] = tab
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gamma_table.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
