"""Host checks of the numeric contract used by the device code."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_scale_char_reciprocal_form_is_exact_on_every_float():
    """sa::scale_char (one double multiply) == ScaleQuantumToChar's correctly
    rounded float division on all 1.2e9 floats in [0, 65535]."""
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "numeric_check")
        subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-pthread",
                        os.path.join(HERE, "native", "numeric_check.cpp"), "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    assert "mismatches 0" in out, out


def test_weights_hex(oracle):
    """16^(-i/15) with the host libm, exactly the reference's doubles."""
    want = ["3ff0000000000000", "3fea9980376421bc", "3fe61c50ae161fd6", "3fe2611186bae675",
            "3fde8e0ffab4f93f", "3fd965fea53d6e3e", "3fd51cb453b9536e", "3fd18c98599396ab",
            "3fcd2cd4a3ec542f", "3fc8406003b2ae5e", "3fc428a2f98d728c", "3fc0c1b77b6cebee",
            "3fbbdb8cdadbe121", "3fb72803c9b50335", "3fb33f972e23dede", "3fb0000000000000"]
    w = oracle.error_weights()
    assert [f"{v.view(np.uint64):016x}" for v in w] == want
    # Closed form within an ulp, and the endpoints exact.
    assert w[0] == 1.0 and w[15] == 1.0 / 16.0
    closed = 16.0 ** (-np.arange(16) / 15.0)
    assert np.max(np.abs(w - closed) / closed) < 1e-15
    # The reference's rounded table (docs/DESIGN.md:1151-1176).
    for i, v in [(0, 1.0), (1, 0.831238), (2, 0.690956), (8, 0.227931), (14, 0.075189), (15, 0.0625)]:
        assert abs(w[i] - v) < 5e-7


def test_scaled_weights_order(oracle):
    """(1/16 * diffusion) * w[k], left to right."""
    w = oracle.error_weights()
    for diff in (1.0, 0.5, 0.75):
        sw = oracle.scaled_weights(diff)
        assert np.array_equal(sw, (np.float64(1 / 16) * np.float64(diff)) * w)


def test_qsort_with_quantize_error_compare_sorts_ascending():
    """The device build replaces QuantizeErrorFlatten + qsort (src/rd_octree.cpp:388-398)
    by a radix selection of errors[n - aggressive].  That is only valid if this
    libc's qsort, with QuantizeErrorCompare (src/rd_octree.cpp:23-29, which returns
    0 for values within MagickEpsilon), still yields the plain ascending order.
    Checked here on arrays full of near-ties, on both of glibc's buffer paths
    (< 1024 bytes on the stack, larger ones malloc'd)."""
    import ctypes as C
    libc = C.CDLL(None)
    cmp_t = C.CFUNCTYPE(C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double))

    def qe_cmp(a, b):
        p, q = a[0], b[0]
        if p > q:
            return 1
        if abs(q - p) <= 1e-12:
            return 0
        return -1

    cmp = cmp_t(qe_cmp)
    libc.qsort.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, cmp_t]
    rng = np.random.default_rng(7)
    for n in (5, 64, 127, 129, 1000, 5000):
        for trial in range(4):
            base = rng.choice([0.0, 1e-6, 0.5, 3.0, 1234.5], n)
            # Clusters of distinct values closer than MagickEpsilon.
            v = base + rng.integers(-3, 4, n) * np.spacing(base + 1e-300) * rng.integers(1, 5, n)
            v = np.abs(v)
            a = np.ascontiguousarray(v.copy())
            libc.qsort(a.ctypes.data, n, 8, cmp)
            assert np.array_equal(a.view(np.uint64), np.sort(v).view(np.uint64)), (n, trial)


def test_cube_midpoint_closed_form_equals_the_descent():
    """The cube midpoint the classification forms level by level
    (src/rd_octree.cpp:318-329: mid = 32767.5, bisect = 32768, then per level
    bisect *= 0.5 and mid +-= bisect) equals (B + 1/2) 2^(16-l) - 1/2 for the
    channel's top l bits B, bit for bit, at every level and every B (the
    device's cell placement uses the closed form)."""
    for level in range(1, 9):
        for b in range(1 << level):
            mid, bisect = 65535.0 / 2.0, (65535.0 + 1.0) / 2.0
            for k in range(1, level + 1):
                bisect *= 0.5
                mid += bisect if (b >> (level - k)) & 1 else -bisect
            closed = (float(b) + 0.5) * float(1 << (16 - level)) - 0.5
            assert mid == closed, (level, b)
