"""Golden digest of the Q1 hex 128^3, kappa = 1 stiffness matrix (no Dirichlet
data), written by the oracle: the pruned pattern and the values of
Builder::buildLhs({AssemblyStiffness(1)}) + closeMatrix() on buildCube<Hex>
128^3 (reference include/builder.hpp:31-154, 366-371; src/mesh.cpp:456-535).

The matrix (44.5 M entries) is too large to commit, and the oracle needs
about 20 s and 4 GB for it, so only sizes and SHA-256 digests of the CSR
arrays are kept: row pointers as int64, columns as int32, values as the
little-endian float64 bits.  The GPU test compares its own arrays' digests.

    python tests/golden/make_c2_golden.py      # rewrites c2_q1hex128.json
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main(n=128):
    t = time.time()
    o = O.Space("hex_q1", "cube_hex", (n, n, n))
    rp, col, v, _ = o.assemble([("stiffness", 1.0)])
    out = {
        "workload": f"Q1 hex {n}^3 unit cube, 2x2x2 Gauss, AssemblyStiffness(1), no Dirichlet, prune(1, 1e-16)",
        "n": n,
        "ndofs": int(len(rp) - 1),
        "nnz": int(len(v)),
        "rowptr_int64_sha256": digest(rp.astype(np.int64)),
        "col_int32_sha256": digest(col.astype(np.int32)),
        "val_float64_sha256": digest(v.astype("<f8")),
        "val_abs_sum": float(np.abs(v).sum()),
    }
    with open(os.path.join(HERE, f"c2_q1hex{n}.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out, indent=1), f"({time.time() - t:.1f} s)")


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 128)
