"""Write tests/golden/reference_known_answers.json.

The data are the known answers that formoniq's own tests assert (luiswirth/
formoniq, paths relative to crates/), transcribed as numbers: hand tables and
closed forms, each recorded with the test it comes from. Nothing here imports
or runs the reference; re-running this script rewrites the same file.
"""
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def unit_laplacian(d):
    # formoniq/tests/unit_elmat.rs:36-49: vol * [[d, -1, ...], [-1, 1, 0..], ...]
    n = d + 1
    m = [[0.0] * n for _ in range(n)]
    m[0][0] = float(d)
    for i in range(1, n):
        m[i][0] = m[0][i] = -1.0
        m[i][i] = 1.0
    vol = 1.0 / math.factorial(d)
    return [[v * vol for v in row] for row in m]


def main():
    data = {
        "laplacian_refcell": {
            "source": "formoniq/tests/unit_elmat.rs:31-49 (dif_both(d, 1) on the reference cell)",
            "cases": {str(d): unit_laplacian(d) for d in range(1, 5)},
        },
        "mass_refcell": {
            "source": "formoniq/tests/unit_elmat.rs:51-79 (mass(d, 0) on the reference cell)",
            "cases": {
                "1": [[1 / 3, 1 / 6], [1 / 6, 1 / 3]],
                "2": [[1 / 12, 1 / 24, 1 / 24], [1 / 24, 1 / 12, 1 / 24], [1 / 24, 1 / 24, 1 / 12]],
                "3": [[1 / 60, 1 / 120, 1 / 120, 1 / 120], [1 / 120, 1 / 60, 1 / 120, 1 / 120],
                      [1 / 120, 1 / 120, 1 / 60, 1 / 120], [1 / 120, 1 / 120, 1 / 120, 1 / 60]],
            },
        },
        "hodge_mass_dim2_grade1": {
            "source": "formoniq/src/operators.rs:931-943",
            "value": [[1 / 3, 1 / 6, 0.0], [1 / 6, 1 / 3, 0.0], [0.0, 0.0, 1 / 6]],
        },
        "dif_trial_n2_k1": {
            "source": "formoniq/src/operators.rs:945-960",
            "value": [[-1 / 2, 1 / 3, 1 / 6], [-1 / 2, 1 / 6, 1 / 3], [0.0, -1 / 6, 1 / 6]],
        },
        "dif_test_n2_k1": {
            "source": "formoniq/src/operators.rs:962-975",
            "value": [[-1 / 2, -1 / 2, 0.0], [1 / 3, 1 / 6, -1 / 6], [1 / 6, 1 / 3, 1 / 6]],
        },
        "kuhn_unit_cube": {
            "source": "simplicial/src/mesher/grid.rs:115-138; regge/src/mesher/cartesian.rs:242-269",
            "cells": [[0, 1, 3, 7], [0, 2, 3, 7], [0, 1, 5, 7], [0, 4, 5, 7], [0, 2, 6, 7], [0, 4, 6, 7]],
            "coords": [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]],
        },
        "kuhn_unit_square_2": {
            "source": "regge/src/mesher/cartesian.rs:271-301",
            "cells": [[0, 1, 4], [0, 3, 4], [1, 2, 5], [1, 4, 5], [3, 4, 7], [3, 6, 7], [4, 5, 8], [4, 7, 8]],
            "coords": [[0.0, 0.0], [0.5, 0.0], [1.0, 0.0], [0.0, 0.5], [0.5, 0.5], [1.0, 0.5],
                       [0.0, 1.0], [0.5, 1.0], [1.0, 1.0]],
        },
        "fdm_laplace_interior": {
            "source": "formoniq/tests/fdm.rs:111-139 (LAPLACE_MATRICES_INTERIOR)",
            "cases": {
                "1": [[2, -1], [-1, 2]],
                "2": [[4, -1, -1, 0], [-1, 4, 0, -1], [-1, 0, 4, -1], [0, -1, -1, 4]],
                "3": [[6, -1, -1, 0, -1, 0, 0, 0], [-1, 6, 0, -1, 0, -1, 0, 0],
                      [-1, 0, 6, -1, 0, 0, -1, 0], [0, -1, -1, 6, 0, 0, 0, -1],
                      [-1, 0, 0, 0, 6, -1, -1, 0], [0, -1, 0, 0, -1, 6, 0, -1],
                      [0, 0, -1, 0, -1, 0, 6, -1], [0, 0, 0, -1, 0, -1, -1, 6]],
            },
        },
        "fem3d_tets": {
            "source": "formoniq/tests/fem3d.rs:40-47 (Kuhn tets of a box, local vertex ids)",
            "tets": [[0, 1, 3, 7], [0, 1, 5, 7], [0, 2, 3, 7], [0, 2, 6, 7], [0, 4, 5, 7], [0, 4, 6, 7]],
        },
    }
    with open(os.path.join(HERE, "reference_known_answers.json"), "w") as f:
        json.dump(data, f, indent=1)


if __name__ == "__main__":
    main()
