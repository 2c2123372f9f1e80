"""Writes the committed golden vectors of the LSCM solve path (tests/golden/*.npz, pins.json).

Run from the repository root after building the oracle:  python tests/golden/make_golden.py

Every expected SOLUTION here comes from a solver independent of the oracle's sparse LU:
scipy's SuperLU (`scipy.sparse.linalg.splu`) and dense `numpy.linalg.solve`.  The assembled systems
and pin sets come from the oracle's restatement of lscm.rs:512-676 (the reference ships no golden
vectors; tests/parity.rs holds only counts and quality bands) and are kept as regression vectors
with SHA-256 checksums, so any later change to indexing, pins or weights shows up as a diff.
"""
import hashlib
import json
import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode())
        h.update(str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()


def chart_fixture(pos, faces, angle, chart):
    ch = oracle.Charts(pos, faces, angle)
    s = ch.system(chart)
    n = len(s["row_ptr"]) - 1
    A = sp.csr_matrix((s["vals"], s["cols"], s["row_ptr"]), shape=(n, n))
    b = s["rhs"][:n]
    x = spla.splu(A.tocsc(), permc_spec="COLAMD").solve(b)
    return dict(vals=s["vals"], cols=s["cols"], row_ptr=s["row_ptr"], rhs=s["rhs"], pin_ids=s["pin_ids"],
                pin_uv=s["pin_uv"], free_ids=s["free_ids"], p2=s["p2"], x=x)


def random_laplacians(count, seed):
    """Tiny SPD systems: weighted graph Laplacians of random connected graphs plus a positive diagonal
    on a few 'pinned' rows, the shape of an LSCM free block.  Answers by dense numpy.linalg.solve."""
    rng = np.random.default_rng(seed)
    out = {}
    for k in range(count):
        n = int(rng.integers(3, 40))
        W = np.zeros((n, n))
        for i in range(1, n):  # a random spanning tree keeps it connected
            j = int(rng.integers(0, i))
            W[i, j] = W[j, i] = rng.uniform(0.1, 3.0)
        for _ in range(n):
            i, j = (int(v) for v in rng.integers(0, n, 2))
            if i != j:
                W[i, j] = W[j, i] = rng.uniform(0.1, 3.0)
        A = np.diag(W.sum(1)) - W
        pinned = rng.choice(n, size=max(1, n // 6), replace=False)
        A[pinned, pinned] += rng.uniform(0.5, 2.0, size=len(pinned))
        b = rng.normal(size=n)
        x = np.linalg.solve(A, b)
        S = sp.csr_matrix(A)
        S.sort_indices()
        out[f"m{k}_vals"] = S.data
        out[f"m{k}_cols"] = S.indices.astype(np.int32)
        out[f"m{k}_row_ptr"] = S.indptr.astype(np.int32)
        out[f"m{k}_b"] = b
        out[f"m{k}_x"] = x
    out["count"] = np.array(count)
    return out


def main():
    oracle.build()
    pins = {}

    pos, faces = oracle.cube(6)
    pw, fw = oracle.weld(pos, faces, 1e-12)
    f = chart_fixture(pw, fw, 30.0, 0)
    np.savez_compressed(os.path.join(HERE, "cube6_chart0.npz"), **f)
    pins["cube6_chart0"] = dict(npins=int(len(f["pin_ids"])), sha256=sha(f["pin_ids"], f["pin_uv"]),
                                system_sha256=sha(f["vals"], f["cols"], f["row_ptr"], f["rhs"]))

    pos, faces = oracle.grid(32, 32)
    pw, fw = oracle.weld(pos, faces, 1e-12)
    f = chart_fixture(pw, fw, 30.0, 0)
    np.savez_compressed(os.path.join(HERE, "grid32_chart0.npz"), **f)
    pins["grid32_chart0"] = dict(npins=int(len(f["pin_ids"])), sha256=sha(f["pin_ids"], f["pin_uv"]),
                                 system_sha256=sha(f["vals"], f["cols"], f["row_ptr"], f["rhs"]))

    np.savez_compressed(os.path.join(HERE, "spd_small.npz"), **random_laplacians(8, 2024))

    with open(os.path.join(HERE, "pins.json"), "w") as fh:
        json.dump(pins, fh, indent=1, sort_keys=True)
        fh.write("\n")
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
