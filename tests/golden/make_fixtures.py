"""Copy the reference's own test fixtures (data files only) into tests/golden/.

The fixtures are the input series and expected outputs that the reference's
test tree holds under tests/good_files/text/ (GPL-3.0, OpenDBA authors). The
expected KAT values of tests/openDBA_test.cu:66-299 are restated as data in
kat_expected.json. Run from the repository root:

    python tests/golden/make_fixtures.py /root/reference
"""
import json
import os
import shutil
import sys

FILES = {
    "test1": ["seq1_test.txt", "seq2_test.txt", "seq1_test_znorm.txt", "seq2_test_znorm.txt",
              "output_results.txt", "output_results_znorm.txt", "output_results_simple.txt",
              "simple1_test.txt", "simple2_test.txt", "simple3_test.txt", "simple4_test.txt"],
    "test2": ["random_short1.txt", "random_short2.txt", "random_short3.txt", "random_trun_short2.txt",
              "random_large0.txt", "random_large1.txt"],
}

# Expected values of the Catch2 KATs (tests/openDBA_test.cu), compared after
# round_to_three (tests/test_utils.cuh:23-36). Index -> value; "exact" entries
# are compared without rounding.
KATS = {
    "same": {"files": ["random_short1.txt", "random_short1.txt"], "open_end": 0, "n": 10,
             "exact": {"0": 1.0},
             "round3": {"1": 0.895, "2": 0.8, "3": 0.644, "4": 0.586, "5": 0.488, "6": 0.357,
                        "7": 0.287, "8": 0.139, "9": 0.017},
             "ref": "tests/openDBA_test.cu:66-107"},
    "different": {"files": ["random_short1.txt", "random_short2.txt"], "open_end": 0, "n": 10,
                  "exact": {"0": 1.0},
                  "round3": {"1": 0.895, "2": 0.796, "3": 0.665, "4": 0.59, "5": 0.456, "6": 0.347,
                             "7": 0.265, "8": 0.137, "9": 0.05},
                  "ref": "tests/openDBA_test.cu:109-150"},
    "three": {"files": ["random_short1.txt", "random_short2.txt", "random_short3.txt"], "open_end": 0, "n": 10,
              "exact": {},
              "round3": {"0": 0.972, "1": 0.901, "2": 0.814, "3": 0.681, "4": 0.56, "5": 0.455, "6": 0.362,
                         "7": 0.244, "8": 0.128, "9": 0.056},
              "ref": "tests/openDBA_test.cu:152-194"},
    "truncated": {"files": ["random_short1.txt", "random_trun_short2.txt"], "open_end": 1, "n": 10,
                  "exact": {"0": 1.0},
                  "round3": {"1": 0.895, "2": 0.796, "3": 0.665, "4": 0.59, "5": 0.488, "6": 0.357,
                             "7": 0.287, "8": 0.139, "9": 0.017},
                  "ref": "tests/openDBA_test.cu:196-245"},
    "large": {"files": ["random_large0.txt", "random_large1.txt"], "open_end": 0, "n": 2048,
              "exact": {"0": 1.0},
              "round3": {"1": 0.997, "2": 1.0, "3": 0.959, "2047": -0.457},
              "ref": "tests/openDBA_test.cu:247-299"},
}


def main(ref_root):
    here = os.path.dirname(os.path.abspath(__file__))
    for sub, names in FILES.items():
        os.makedirs(os.path.join(here, sub), exist_ok=True)
        for nm in names:
            shutil.copyfile(os.path.join(ref_root, "tests", "good_files", "text", sub, nm),
                            os.path.join(here, sub, nm))
    with open(os.path.join(here, "kat_expected.json"), "w") as f:
        json.dump(KATS, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
