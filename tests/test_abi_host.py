"""CPU tests of the C ABI and the host-side pieces of libmqw (no GPU).

The library must load and export every entry point include/mqw.h declares;
host pieces (unpack, schedule, Monte-Carlo times, Chebyshev coefficients, SK
generator, file surface) are checked against the oracle and golden vectors.
Reference: MultiQW.cpp."""
import ctypes as C
import hashlib
import os
import re

import numpy as np
import pytest

import multistageqw_amd as M
from multistageqw_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def _declared():
    text = open(os.path.join(ROOT, "include", "mqw.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mqw_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = M.load()
    declared = _declared()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(L.EXPORTS) == declared


def test_abi_version_and_defaults():
    lib = M.load()
    assert lib.mqw_abi_version() == L.ABI_VERSION
    o = M.make_opts()
    assert (o.samples, o.tol, o.precision, o.energy_mode, o.use_graph) == (100, 1e-6, 0, 0, 1)


def test_create_reports_status_without_crashing():
    lib = M.load()
    h = C.c_void_p()
    st = lib.mqw_create(0, C.byref(h))
    assert st in (L.OK, L.EDEVICE)
    if st == L.OK:
        lib.mqw_destroy(h)
    assert lib.mqw_create(-1, C.byref(h)) in (L.EINVAL, L.EDEVICE)
    assert lib.mqw_status_string(L.EIO) == b"i/o error"


def test_schedule_and_times_match_oracle(oracle):
    for n in (4, 9, 16, 20, 28):
        for m in (1, 2, 3, 5, 8):
            rec = M.generate_sk(n, 1, 31 * n + m)[0]
            s = M.schedule_heuristic(n, m, rec)
            J, h = oracle.unpack(n, rec)
            so = oracle.schedule(n, m, J, h)
            assert np.array_equal(s.gamma, so["gamma"]) and np.array_equal(s.short_t, so["short_t"])
            assert s.total_t == so["total_t"] and s.hp2 == so["hp2"] and s.delta2 == so["delta2"]
            assert np.array_equal(M.sample_times(m, 100, s.short_t), oracle.sample_times(m, 100, so["short_t"]))


def test_times_match_libstdcxx_golden():
    g = np.load(os.path.join(GOLD, "mt19937_uniform.npz"))
    assert np.array_equal(M.sample_times(g["s"].size, 100, g["s"]), g["draws_native"])


def test_unpack_matches_oracle(oracle):
    for n in (4, 7, 30):
        rec = np.random.default_rng(n).standard_normal(n * (n + 1) // 2)
        J, h = M.unpack(n, rec)
        Jo, ho = oracle.unpack(n, rec)
        assert np.array_equal(J, Jo) and np.array_equal(h, ho)


def test_chebyshev_coeffs_vs_scipy_and_oracle(oracle):
    g = np.load(os.path.join(GOLD, "scipy_special.npz"))
    for si, s in enumerate(g["s"]):
        c = M.chebyshev_coeffs(s)
        k = np.arange(c.size)
        ref = np.where(k == 0, 1.0, 2.0) * np.where((k >> 1) & 1, -1.0, 1.0) * g["jv"][si, :c.size]
        assert np.max(np.abs(c - ref)) < 1e-13
        co = oracle.cheb_coeffs(s)
        assert c.size == co.size and np.max(np.abs(c - co)) < 1e-15
    assert M.chebyshev_coeffs(0.0).tolist() == [1.0]
    with pytest.raises(M.MQWError):
        M.chebyshev_coeffs(float("nan"))


def test_sk_generator_fixture_and_statistics():
    recs = M.generate_sk(10, 64, 1)
    with open(os.path.join(GOLD, "sk_n10_m2_seed1.bin"), "rb") as f:
        assert recs.astype("<f8").tobytes() == f.read()
    assert hashlib.sha256(recs.astype("<f8").tobytes()).hexdigest() == \
        "71fc310eadbf5dbb2582afd131aeaa37d9792dfee07c69b352cc316cfa5c34fa"
    big = M.generate_sk(20, 500, 9).ravel()
    assert abs(big.mean()) < 0.01 and abs(big.std() - 1) < 0.01
    assert not np.array_equal(M.generate_sk(10, 2, 1), M.generate_sk(10, 2, 2))


def test_file_surface(tmp_path):
    path = os.path.join(GOLD, "sk_n10_m2_seed1.bin")
    assert M.count_instances(path, 10) == 64
    r = M.read_instances(path, 10, 5, 3)
    assert np.array_equal(r, np.fromfile(path, "<f8").reshape(64, 55)[5:8])
    with pytest.raises(M.MQWError):
        M.read_instances(path, 10, 60, 10)           # short file is an error, not a repeat
    with pytest.raises(M.MQWError):
        M.read_instances(str(tmp_path / "missing"), 10, 0, 1)
    M.write_output(10, 2, np.array([0.5, 0.25], np.float32), str(tmp_path))
    M.write_output(10, 2, np.array([0.125], np.float32), str(tmp_path))   # append mode (MultiQW.cpp:266)
    assert np.fromfile(str(tmp_path / "output_10_2"), "<f4").tolist() == [0.5, 0.25, 0.125]


def test_invalid_arguments_are_errors():
    lib = M.load()
    rec = np.zeros(55)
    s = L.Schedule()
    assert lib.mqw_schedule_heuristic(3, 2, L.f64p(np.zeros(6)), C.byref(s)) == L.EINVAL   # n < 4
    assert lib.mqw_schedule_heuristic(10, 0, L.f64p(rec), C.byref(s)) == L.EINVAL         # m < 1
    bad = rec.copy()
    bad[3] = np.nan
    assert lib.mqw_schedule_heuristic(10, 2, L.f64p(bad), C.byref(s)) == L.EINVAL        # NaN parameter
    assert lib.mqw_sample_times(1, 0, L.f32p(np.ones(1, np.float32)), L.f32p(np.ones(1, np.float32))) == L.EINVAL


def test_format_line_matches_iostream_defaults():
    assert M.format_line(np.float32(0.55484647), np.float32(0.25346103)) == "0.554846 0.253461"
    assert M.format_line(np.float32(12.0), np.float32(1.5e-5)) == "12 1.5e-05"


def _append_block(args):
    d, k, size = args
    vals = np.full(size, float(k), np.float32)
    return M.write_output(10, 2, vals, directory=d)


def test_concurrent_appends_keep_blocks_whole(tmp_path):
    """Slices of one run append their blocks concurrently (run.sh:27-38); each lands whole."""
    import multiprocessing as mp
    size, procs = 1 << 20, 6  # 4 MiB per block: many write() calls' worth
    with mp.get_context("spawn").Pool(procs) as pool:
        pool.map(_append_block, [(str(tmp_path), k, size) for k in range(procs)])
    out = np.fromfile(str(tmp_path / "output_10_2"), "<f4")
    assert out.size == procs * size
    blocks = out.reshape(procs, size)
    assert sorted(int(b[0]) for b in blocks) == list(range(procs))
    for b in blocks:
        assert np.all(b == b[0])
