"""Writes int_hash_vectors.json: outputs of the reference's own integer-hash
RNG (src/pscbfm/rngs/Hash.h, first rng32() for a seed and a subsequence), as
built from its source into oracle/_ref/hash_ref by oracle/ref.mk.

Run in a tree where oracle/_ref/hash_ref exists:
    python tests/golden/make_int_hash_vectors.py
"""
import json
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
EXE = HERE.parent.parent / "oracle" / "_ref" / "hash_ref"


def main() -> None:
    rng = np.random.default_rng(20261203)
    pairs = [(0, 0), (1, 0), (0, 1), (0xFFFFFFFF, 0xFFFFFFFF), (0xDEADBEEF, 7)]
    pairs += [(int(a), int(b)) for a, b in zip(rng.integers(0, 2**32, 59), rng.integers(0, 2**28, 59))]
    text = "".join(f"{a} {b}\n" for a, b in pairs)
    out = subprocess.run([str(EXE)], input=text, capture_output=True, text=True, check=True).stdout.split()
    vec = [{"seed": a, "subsequence": b, "rng32": int(v)} for (a, b), v in zip(pairs, out)]
    (HERE / "int_hash_vectors.json").write_text(json.dumps({"source": "src/pscbfm/rngs/Hash.h:39-44,64-68",
                                                           "vectors": vec}, indent=1) + "\n")


if __name__ == "__main__":
    main()
