"""Fixtures from the input data the reference ships (data/AdamData, data/TimData).

The reference holds SK instance files (data/AdamData/description.txt: n(n+1)/2
float64 per instance, lower triangle of J row by row, then h) and no expected
outputs.  This script, run in a container where the reference tree is at
REF (default /root/reference), writes:

  * ref_<set>_sk<n>_first<k>.bin : the first k instances of a few files, as
    committed inputs (byte copies of the data, for tests that run where the
    reference tree is absent);
  * ref_argmin.npz : for every instance of every file, the ground-state index of
    the reference's energy walk (CPU restatement, MultiQW.cpp:325-358) and, for
    the first EXACT_MAX instances of each file, whether it equals the argmin of
    exact FP64 energies (sum_{i<j} J_ij s_i s_j + sum_i h_i s_i, brute force);
    and the walk's E_0 for the instances of the committed inputs.

Run from the repository root: python tests/golden/make_reference_data.py
"""
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

REF = os.environ.get("REF", "/root/reference")
EXACT_MAX = 200
SUBSETS = [("adam", 10, 64), ("adam", 16, 16), ("adam", 20, 8), ("tim", 8, 32), ("tim", 18, 8)]
DIRS = {"adam": "AdamData", "tim": "TimData"}


def records(path, n):
    P = n * (n + 1) // 2
    raw = np.fromfile(path, "<f8")
    return raw[: (raw.size // P) * P].reshape(-1, P)


def compat(n, rec):
    J, h = O.unpack(n, rec)
    _, _, info = O.energy_compat(n, J, h)
    return info["eloc"], info["e0"]


def exact_argmin(n, recs):
    N = 1 << n
    k = np.arange(N, dtype=np.uint32)
    S = np.stack([np.where((k >> i) & 1, 1.0, -1.0) for i in range(n)], axis=1)
    out = []
    for rec in recs:
        J, h = O.unpack(n, rec)
        Jd, hd = J.astype(np.float64), h.astype(np.float64)
        E = S @ hd + np.einsum("ki,ki->k", S @ Jd, S)
        out.append(int(np.argmin(E)))
    return np.array(out, np.uint32)


def main():
    O.lib()
    table = {}
    for key, d in DIRS.items():
        for fn in sorted(os.listdir(os.path.join(REF, "data", d))):
            if not fn.startswith("SK_") or not fn.endswith("n"):
                continue
            n = int(fn[3:-1])
            recs = records(os.path.join(REF, "data", d, fn), n)
            with ThreadPoolExecutor(8) as ex:
                res = list(ex.map(lambda r: compat(n, r), recs))
            eloc = np.array([r[0] for r in res], np.uint32)
            e0 = np.array([r[1] for r in res], np.float64)
            ex_n = min(len(recs), EXACT_MAX)
            agree = exact_argmin(n, recs[:ex_n]) == eloc[:ex_n]
            table[f"{key}_{n}_eloc"] = eloc
            table[f"{key}_{n}_exact_agree"] = agree
            for sk, sn, cnt in SUBSETS:
                if sk == key and sn == n:
                    recs[:cnt].astype("<f8").tofile(os.path.join(HERE, f"ref_{key}_sk{n}_first{cnt}.bin"))
                    table[f"{key}_{n}_e0_first{cnt}"] = e0[:cnt]
            print(f"{d}/{fn}: {len(recs)} instances, compat vs exact argmin agree on {int(agree.sum())}/{ex_n}",
                  flush=True)
    np.savez_compressed(os.path.join(HERE, "ref_argmin.npz"), **table)


if __name__ == "__main__":
    main()
