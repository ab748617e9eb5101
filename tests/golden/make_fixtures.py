"""Regenerate the committed golden fixtures of tests/golden/.

Sources are independent of the code under test:
  * mt19937_uniform.npz - libstdc++'s std::mt19937 / uniform_real_distribution<float>,
    built with g++ -O3 -march=native (the reference's build line, README.md:20) and
    with plain -O3 for contrast (mt19937_uniform.cpp);
  * scipy_special.npz   - scipy.special.jv (Bessel J_k) and scipy.special.ndtri;
  * sk_n10_m2_seed1.bin - the 64-instance n=10 input file (55 float64 per instance),
    written by the package's seeded SK generator; its SHA-256 is pinned in the tests.
Run from the repository root: python tests/golden/make_fixtures.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
import scipy.special as sp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
S_VALUES = ["0.1037676", "0.45107886", "1.0", "3.3333", "0.05", "7.77", "12.5"]


def mt_vectors():
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for tag, flags in [("native", ["-O3", "-march=native"]), ("plain", ["-O3"])]:
            exe = os.path.join(d, f"mt_{tag}")
            subprocess.run(["g++", *flags, os.path.join(HERE, "mt19937_uniform.cpp"), "-o", exe], check=True)
            raw = subprocess.run([exe, *S_VALUES], check=True, stdout=subprocess.PIPE).stdout
            out[f"raw_{tag}"] = np.frombuffer(raw[:4000], np.uint32)
            out[f"draws_{tag}"] = np.frombuffer(raw[4000:], np.float32).reshape(len(S_VALUES), 100)
    assert np.array_equal(out["raw_native"], out["raw_plain"])
    np.savez(os.path.join(HERE, "mt19937_uniform.npz"), s=np.array(S_VALUES, np.float32),
             raw=out["raw_native"], draws_native=out["draws_native"], draws_plain=out["draws_plain"])


def scipy_vectors():
    s = np.array([1e-4, 0.05, 0.7, 2.5, 9.0, 31.4, 64.0, 150.0, 480.0, 1200.0])
    k = np.arange(0, 1400)
    jv = sp.jv(k[None, :], s[:, None])
    p = np.concatenate([np.array([2.0 ** -e for e in range(1, 31)], np.float32).astype(np.float64),
                        np.array([1 / (np.float32(2.718281828459045) * np.float32(2.0 ** e)) for e in range(1, 31)],
                                 np.float32).astype(np.float64),
                        np.linspace(0.001, 0.999, 41)])
    np.savez(os.path.join(HERE, "scipy_special.npz"), s=s, k=k, jv=jv, p=p, ndtri=sp.ndtri(p))


def c1_input():
    sys.path.insert(0, ROOT)
    import multistageqw_amd as M
    recs = M.generate_sk(10, 64, 1)
    recs.astype("<f8").tofile(os.path.join(HERE, "sk_n10_m2_seed1.bin"))


if __name__ == "__main__":
    mt_vectors()
    scipy_vectors()
    c1_input()
    print("fixtures written to", HERE)
