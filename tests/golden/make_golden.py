"""Write the known-answer fixtures of the reference's own tests as JSON.

The values are transcribed from the reference's test sources (data only):
  friction_contact_test.cu:10-21   exact_ee_uv on two crossing edges
  friction_contact_test.cu:23-93   solve_coulomb_contact cases and the
                                   complementarity / cone sweep
  broadphase/main.cpp:353-380      touching boxes must pair
  interval_root_finder.cpp:516-517 CCD numerical filters (VF 34*2^-23,
                                   EE 32*2^-23, derive_numerical_filter.py)

Run:  python tests/golden/make_golden.py   (rewrites the .json next to it)
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

fixtures = {
    "ee_uv_scale_independent": {
        "x": [[-0.005, 0.0, 0.0], [0.005, 0.0, 0.0],
              [0.0, -0.005, 0.0005], [0.0, 0.005, 0.0005]],
        "eps": 1e-6,
        "expected_uv": [0.5, 0.5],
    },
    "coulomb_exact": [
        {"name": "separating_unchanged", "u": [1.0, -2.0, 0.5], "n": [0, 0, 1],
         "mu": 0.7, "expected": [1.0, -2.0, 0.5]},
        {"name": "stick_inside_cone", "u": [0.2, 0.0, -1.0], "n": [0, 0, 1],
         "mu": 0.5, "expected": [0.0, 0.0, 0.0]},
        {"name": "slide_on_cone_boundary", "u": [2.0, 0.0, -1.0], "n": [0, 0, 1],
         "mu": 0.5, "expected": [1.5, 0.0, 0.0]},
        {"name": "frictionless_removes_penetration", "u": [2.0, -3.0, -1.0],
         "n": [0, 0, 1], "mu": 0.0, "expected": [2.0, -3.0, 0.0]},
    ],
    "coulomb_sweep": {
        "n": [0, 0, 1],
        "u": [[0.0, 0.0, 1.0], [2.0, -1.0, 0.0], [0.1, 0.2, -1.0],
              [2.0, 0.0, -1.0], [-3.0, 4.0, -0.5], [0.0, 0.0, -2.0]],
        "mu": [0.0, 0.1, 0.5, 2.0],
        "tolerance": 1e-5,
    },
    # one self query over two edge boxes that share the face x = 1
    "touching_boxes": {
        "boxes": [[0.0, 0.0, 0.0, 1.0, 1.0, 1.0],
                  [1.0, 0.25, 0.25, 2.0, 0.75, 0.75]],
        "expected_pairs": [[0, 1]],
    },
    "ccd_filters": {
        "vf": 34 * 2.0 ** -23,
        "ee": 32 * 2.0 ** -23,
        "vf_literal": 4.053116e-06,
        "ee_literal": 3.814698e-06,
    },
}

if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(fixtures, f, indent=1)
    print("wrote reference_kats.json")
