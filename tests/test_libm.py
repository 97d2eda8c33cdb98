"""The device's exp / pow / log (trinity_amd/csrc/trn_libm.hpp) against the
host C library the reference links, on the CPU.

The header is compiled with g++ into tests/libm_harness.cpp; the same source
is what the GPU kernels inline.  Bars:
  exp, pow   bit-identical to the host libm on every input drawn (millions
  and cube   over the ranges the path uses, the scaled and subnormal ranges,
             special values);
  log        correctly rounded (checked against 300-bit mpmath values), so
             equal to the host's log wherever the host's is correctly
             rounded: no difference on metric eigenvalues (>= 6.25), one-ulp
             differences on at most 0.3 % of arguments near 1, where the host
             log itself is off by one ulp.
  log1p_small, expm1 (the adjoint's divided differences; the CPU reference
             adjoint calls the same code): within one ulp of 300-bit values.
"""
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "trinity_amd", "csrc")


@pytest.fixture(scope="module")
def harness():
    d = tempfile.mkdtemp(prefix="trn_libm_")
    exe = os.path.join(d, "libm_harness")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe,
                    os.path.join(ROOT, "tests", "libm_harness.cpp"), "-lm"], check=True)
    return exe


def sweep(exe, op, n, seed):
    out = subprocess.run([exe, "sweep", op, str(n), str(seed)], check=True,
                         capture_output=True, text=True).stdout
    res = {}
    for line in out.splitlines():
        name, bad, tot = line.split()
        res[name] = (int(bad), int(tot))
    return res


def evaluate(exe, op, x, y=None):
    d = tempfile.mkdtemp(prefix="trn_libm_eval_")
    fi, fo = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
    a = np.asarray(x, dtype=np.float64)
    if y is not None:
        a = np.stack([a, np.asarray(y, dtype=np.float64)], axis=1).ravel()
    a.tofile(fi)
    subprocess.run([exe, "eval", op, fi, fo], check=True)
    return np.fromfile(fo, dtype=np.float64)


def same_bits(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    return (a.view(np.uint64) == b.view(np.uint64)) | both_nan


def test_tables_match_their_generator():
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen", os.path.join(CSRC, "gen_libm_tables.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    text = open(os.path.join(CSRC, "trn_libm_tables.inc")).read()
    for t, h in gen.exp_table():
        assert f"0x{t:016x}ULL, 0x{h:016x}ULL" in text
    for a, b, c in gen.log_table():
        assert f"{{{a.hex()}, {b.hex()}, {c.hex()}}}" in text
    # spot values fixed by the mathematics: 2^(1/128) rounds to
    # 0x3ff0163da9fb3335; the sub-interval holding 1.0 has 1/c = 1, log c = 0
    tab = gen.exp_table()
    assert tab[1][1] + (1 << 45) == 0x3FF0163DA9FB3335
    rows = gen.log_table()
    i1 = ((0x3FF0000000000000 - gen.OFF) >> 45) % 128
    assert rows[i1] == (1.0, 0.0, 0.0)


def test_exp_bit_identical_to_host_libm(harness):
    for name, (bad, tot) in sweep(harness, "exp", 1_000_000, 11).items():
        assert bad == 0, f"{name}: {bad} of {tot} differ"


def test_pow_bit_identical_to_host_libm(harness):
    for name, (bad, tot) in sweep(harness, "pow", 500_000, 12).items():
        assert bad == 0, f"{name}: {bad} of {tot} differ"


def test_cube_bit_identical_to_host_pow(harness):
    """libm::cube (the quality's pow(x, 3), mostly from the double-double
    cube) equals the host's pow(x, 3.0) bit for bit, on the quality's
    arguments in [0, 1] and on general, tiny and huge ones."""
    for name, (bad, tot) in sweep(harness, "cube", 1_000_000, 14).items():
        assert bad == 0, f"{name}: {bad} of {tot} differ"


def test_log_matches_host_where_host_is_correctly_rounded(harness):
    res = sweep(harness, "log", 1_000_000, 13)
    assert res["log0"][0] <= 20          # wide range (host misrounds ~2e-6)
    assert res["log3"][0] <= 20          # subnormal arguments
    for name in ("log1", "log2"):        # near 1: host misrounds ~2e-3
        bad, tot = res[name]
        assert bad <= 3e-3 * tot, name


def test_log_correctly_rounded_against_mpmath(harness):
    import mpmath
    mpmath.mp.prec = 300
    rng = np.random.default_rng(3)
    x = np.concatenate([np.exp(rng.uniform(-700, 700, 4000)), 1 + rng.uniform(0, 3, 4000),
                        1 + rng.uniform(-1e-3, 1e-3, 4000), rng.uniform(6.25, 1e6, 4000),
                        1 + rng.uniform(-0.006, 0.006, 4000)])
    got = evaluate(harness, "log", x)
    want = np.array([float(mpmath.log(mpmath.mpf(float(v)))) for v in x])
    assert same_bits(got, want).all()


def test_log_equals_host_on_metric_eigenvalues(harness):
    rng = np.random.default_rng(4)
    # eigenvalues of clamped metrics lie in [1/h_max^2, 1/h_min^2] = [6.25, 1e18]
    x = np.exp(rng.uniform(np.log(6.25), np.log(1e18), 400_000))
    got = evaluate(harness, "log", x)
    assert same_bits(got, [_libm().log(v) for v in x]).all()


def test_special_values_match_host(harness):
    inf, nan = float("inf"), float("nan")
    c = _libm()
    sub = struct.unpack("<d", struct.pack("<Q", 1))[0]
    xs = [0.0, -0.0, 1.0, -1.0, 2.0, -2.0, 0.5, inf, -inf, nan, sub, 1e-310, 1e308, -1e-310,
          709.78, 709.79, -745.13, -745.14, 1e-20, -1e-20, 3.0]
    got_e = evaluate(harness, "exp", xs)
    got_l = evaluate(harness, "log", xs)
    for i, x in enumerate(xs):
        assert same_bits([got_e[i]], [c.exp(x)]).all(), x
        assert same_bits([got_l[i]], [c.log(x)]).all(), x
    ys = [0.0, -0.0, 1.0, -1.0, 2.0, 3.0, -3.0, 0.5, -0.25, inf, -inf, nan, 1e300, -1e300,
          1e-30, 2.5]
    xb, yb, want = [], [], []
    for x in xs:
        for y in ys:
            xb.append(x)
            yb.append(y)
            want.append(c.pow(x, y))
    got = evaluate(harness, "pow", xb, yb)
    ok = same_bits(got, want)
    assert ok.all(), [(xb[i], yb[i], got[i], want[i]) for i in np.flatnonzero(~ok)[:5]]


def _ulps(got, want):
    g = np.asarray(got, dtype=np.float64).view(np.int64)
    w = np.asarray(want, dtype=np.float64).view(np.int64)
    return np.abs(g - w)


def test_adjoint_helpers_within_one_ulp(harness):
    import mpmath
    mpmath.mp.prec = 300
    rng = np.random.default_rng(5)
    r = np.concatenate([rng.uniform(-2.0 ** -10, 2.0 ** -10, 6000),
                        rng.uniform(-1e-4, 1e-4, 6000), [0.0, 1e-300, -1e-300]])
    got = evaluate(harness, "log1p_small", r)
    want = [float(mpmath.log1p(mpmath.mpf(float(v)))) for v in r]
    assert _ulps(got, want).max() <= 1
    d = np.concatenate([rng.uniform(-1 / 16, 1 / 16, 4000), rng.uniform(-0.7, 0.7, 4000),
                        rng.uniform(-40, 40, 4000), rng.uniform(-1e-8, 1e-8, 1000),
                        [0.0, -0.0, 0.0625, -0.0625, 0.7, -0.7]])
    got = evaluate(harness, "expm1", d)
    want = [float(mpmath.expm1(mpmath.mpf(float(v)))) for v in d]
    assert _ulps(got, want).max() <= 1


def _libm():
    """The host C library (the one the reference links), called directly."""
    import ctypes
    lib = ctypes.CDLL("libm.so.6")
    for name, nargs in (("exp", 1), ("log", 1), ("pow", 2)):
        f = getattr(lib, name)
        f.restype = ctypes.c_double
        f.argtypes = [ctypes.c_double] * nargs
    return lib
