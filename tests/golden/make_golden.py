"""Writes tests/golden/golden.npz: known-answer vectors computed with numpy / scipy / pure
Python, independent of the oracle.  The reference ships no fixtures of its own; its README
sample's printed parameters are recorded here as well.

    python tests/golden/make_golden.py
"""
import json
import os

import numpy as np
from scipy.signal import windows

HERE = os.path.dirname(os.path.abspath(__file__))


def reverse_hash_set_semantics(J, n, B, a, ai, thresh, vote, hits, J2, W):
    m = n // B
    S = {(int(j) * ai) % W for j in J2}
    mx = max(S)
    for Jv in J:
        low = (int(Jv) * m - m // 2) % n
        if (low % W) > mx:
            continue
        for d in range(m):
            p = (low + d) % n
            if (p % W) in S:
                loc = (p * a) % n
                vote[loc] += 1
                if vote[loc] == thresh:
                    hits.append(loc)


def main():
    rng = np.random.default_rng(20260101)
    out = {}
    for l2 in (3, 10, 12):
        n = 1 << l2
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"fft_in_{l2}"] = x
        out[f"fft_out_{l2}"] = np.fft.fft(x)
        out[f"ifft_out_{l2}"] = np.fft.ifft(x) * n
    for N in (101, 1849):
        x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        out[f"dft_in_{N}"] = x
        out[f"dft_out_{N}"] = np.fft.fft(x)
        out[f"chebwin_{N}"] = windows.chebwin(N, at=160.0, sym=True)
    # cutoff with ties: value descending, then lower index
    s = rng.integers(0, 6, 4096).astype(np.float64)
    K = 1000
    out["cutoff_keys"] = s
    out["cutoff_K"] = np.array([K])
    out["cutoff_out"] = np.sort(np.lexsort((np.arange(s.size), -s))[:K]).astype(np.int32)
    # reverse hash: three location loops, threshold 2
    n, B, W = 1 << 13, 64, 1 << 10
    J2 = np.sort(rng.choice(W, 200, replace=False)).astype(np.int32)
    vote = np.zeros(n, dtype=np.int64)
    hits = []
    loops = []
    for _ in range(3):
        a = int(rng.integers(0, n // 2)) * 2 + 1
        ai = pow(a, -1, n)
        J = np.sort(rng.choice(B, 24, replace=False)).astype(np.int32)
        reverse_hash_set_semantics(J, n, B, a, ai, 2, vote, hits, J2, W)
        loops.append((a, ai, J))
    out["rh_dims"] = np.array([n, B, W, 2])
    out["rh_J2"] = J2
    out["rh_a"] = np.array([l[0] for l in loops], dtype=np.int32)
    out["rh_ai"] = np.array([l[1] for l in loops], dtype=np.int32)
    out["rh_J"] = np.stack([l[2] for l in loops])
    out["rh_hits"] = np.array(hits, dtype=np.int32)
    out["rh_vote"] = vote.astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    readme = {"args": [27, 16000, 0.65, 2, 2, 7, 4096, 2], "B_thresh": 32000,
              "W_RLFC": 4194304, "numlobes": 183314.0, "b_f": 966, "B": 131072, "w": 2230607,
              "source": "README.md:84-90 (printed by ./exec 27 16000 0.65 2 2 7 4096 2)"}
    with open(os.path.join(HERE, "readme_sample.json"), "w") as fh:
        json.dump(readme, fh, indent=1)


if __name__ == "__main__":
    main()
