"""Write the golden fixtures under tests/golden/ (run in the build container).

rng_kat.json     Known-answer vectors of the reference's rust-random123 tests,
                 read as data from rust-random123/src/philox.rs:242-296 and
                 rust-random123/src/threefry.rs:97-142 (test arrays, keys and
                 the seed_from_u64(42) vector), plus seed -> key values derived
                 with the oracle (pinned by the seed 42 vector).
solvers_kat.json Closed-form cases of src/solvers.rs:290-410 (inputs and
                 expected outputs as data).

Usage: python tests/golden/make_golden.py [/root/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))


def _array(text, name):
    m = re.search(name + r"\s*:\s*\[[^\]]*\]\s*=\s*\[(.*?)\];", text, re.S)
    return [int(t, 16) for t in re.findall(r"0x[0-9a-fA-F]+", m.group(1))]


def _const(text, name):
    m = re.search(r"const\s+" + name + r"\s*:\s*u\d+\s*=\s*(0x[0-9a-fA-F]+)", text)
    return int(m.group(1), 16)


def main(ref):
    from oracle import rng as R

    philox = open(os.path.join(ref, "rust-random123/src/philox.rs")).read()
    threefry = open(os.path.join(ref, "rust-random123/src/threefry.rs")).read()
    seeded = re.search(r"seed_from_u64\(42\);\s*//\s*assert_eq!\(rng\.next_u64\(\),\s*(\d+)\)", threefry)
    kat = {
        "source": "rust-random123/src/philox.rs:242-296, rust-random123/src/threefry.rs:97-142",
        "philox_seed1": _const(philox, "SEED1"),
        "philox_seed2": _const(philox, "SEED2"),
        "philox2x32": _array(philox, "TEST_VEC_2X32"),
        "philox2x64": _array(philox, "TEST_VEC_2X64"),
        "philox4x32": _array(philox, "TEST_VEC_4X32"),
        "philox4x64": _array(philox, "TEST_VEC_4X64"),
        "threefry_key": [_const(threefry, "SEED1_U64"), _const(threefry, "SEED2_U64")],
        "threefry2x64": _array(threefry, "TEST_VEC_1"),
        "threefry_seed42_first_u64": int(seeded.group(1)),
    }
    s0 = R.Stream.from_seed(R.THREEFRY2X64, 0)
    p42 = R.Stream.from_seed(R.PHILOX4X32, 42)
    kat["derived"] = {
        "threefry_seed0_key": list(s0.key),
        "threefry_seed0_words": [int(w) for w in s0.words64(0, 4)],
        "philox_seed42_key": list(p42.key),
        "philox_seed42_block0": [int(x) for x in R.philox4x32([[0, 0, 0, 0]], p42.key)[0]],
        "philox_key42_0_block0": [int(x) for x in R.philox4x32([[0, 0, 0, 0]], (42, 0))[0]],
    }
    with open(os.path.join(HERE, "rng_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)

    solvers = {
        "source": "src/solvers.rs:290-410",
        "upper_basic": {"u": [[2, 1, 1], [0, 2, 1], [0, 0, 2]], "y": [4, 2, 2], "x": [1.25, 0.5, 1.0]},
        "upper_identity": {"u": [[1, 0, 0], [0, 1, 0], [0, 0, 1]], "y": [1, 2, 3], "x": [1, 2, 3]},
        "upper_zero_diag": {"u": [[1, 1, 1], [0, 0, 1], [0, 0, 1]], "y": [1, 1, 1], "x_row2": 1.0},
        "upper_f32": {"u": [[2, 1], [0, 2]], "y": [3, 2], "x": [1, 1]},
        "lsqr_simple": {"a": [[1, 0], [1, 1], [0, 1]], "b": [1, 0, -1], "x": [1, -1], "tol": 1e-2},
    }
    with open(os.path.join(HERE, "solvers_kat.json"), "w") as f:
        json.dump(solvers, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
