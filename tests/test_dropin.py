"""The C++ drop-in API (include/rd_riemersma.h, rd_octree.h, rd_plugin.h).

CPU: the drop-in library exports every engine entry point the headers
declare, and a caller written against the headers compiles and links.
GPU: that caller (tests/native/dropin_test.cpp) runs every engine through the
reference's signatures and compares with the oracle on the same inputs.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "dropin_test.cpp")
OUT = os.path.join(ROOT, "tests", "native", "_build", "dropin_test")
LIB = os.path.join(ROOT, "stipple_amd", "librd_amd.so")
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")

# Engine entry points of the reference API this layer must provide
# (include/rd_riemersma.h:178-183, :204-217, :278-309, :450-481 and
# include/rd_plugin.h:133-148 of the reference).
EXPECTED = [
    "rd::RiemersmaWalkCuda(", "rd::RiemersmaWalkBatch(", "rd::RiemersmaBlocksCuda(",
    "rd::RiemersmaApproxCuda(", "rd::ApproxInverseFilter(", "rd::BuildCurveIndex(",
    "rd::BlockCurveOrder(", "rd::BlockCurveOwner(", "rd::HilbertPoint(", "rd::CudaAvailable()",
    "rd::CudaMaxFrames(", "rd::CudaAllocPinned(", "rd::CudaFreePinned(", "rd::ColorTree::Build(",
    "rd::RegisterDither(", "rd::FindDither(", "rd::RegisteredDithers()", "rd::DitherListText",
    "rd::ImWritePalette(", "rd::ImReadPalette(", "rd::ImBuildPaletteFromPixels(",
]


def _build():
    newest = max(os.path.getmtime(p) for p in (SRC, LIB, ORACLE))
    if os.path.exists(OUT) and os.path.getmtime(OUT) >= newest:
        return OUT
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    subprocess.run(
        ["g++", "-std=c++17", "-O1", "-Wall", "-I" + os.path.join(ROOT, "include"), SRC,
         "-L" + os.path.join(ROOT, "stipple_amd"), "-lrd_amd",
         "-L" + os.path.join(ROOT, "oracle"), "-loracle",
         "-Wl,-rpath,$ORIGIN/../../../stipple_amd", "-Wl,-rpath,$ORIGIN/../../../oracle",
         "-o", OUT], check=True)
    return OUT


def test_dropin_library_exports_the_engine_api():
    assert os.path.exists(LIB), "librd_amd.so not built"
    syms = subprocess.run(["nm", "-D", "-C", "--defined-only", LIB], check=True,
                          capture_output=True, text=True).stdout
    missing = [s for s in EXPECTED if s not in syms]
    assert not missing, missing


def test_dropin_caller_compiles_and_links():
    assert os.path.exists(_build())


@pytest.mark.gpu
def test_dropin_engines_match_the_oracle(gpu):
    exe = _build()
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK: 0 failure(s)" in r.stdout


# ---- the reference-side adapter (stipple_amd/host/rd_hip_engine.cpp) -------------------
REFERENCE = "/root/reference"
ADAPTER = os.path.join(ROOT, "oracle", "_ref", "adapter_test")
# What the reference's CUDA translation units and their stub define for the rest
# of the reference (src/rd_riemersma_cuda.cu:460-554, src/rd_blocks_cuda.cu:1006,
# :1720-1726, src/rd_approx_cuda.cu:232, :386, src/rd_cuda_stub.cpp:50-95).
ADAPTER_DEFINES = ["rd::CudaAvailable()", "rd::CudaMaxFrames(", "rd::RiemersmaWalkCuda(",
                   "rd::RiemersmaBlocksCuda(", "rd::CudaAllocPinned(", "rd::CudaFreePinned(",
                   "rd::RiemersmaApproxCuda(", "rd::ApproxInverseFilter("]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "src")),
                    reason="the reference checkout is not present")
def test_reference_adapter_links_into_the_reference_build():
    """The adapter compiles against the reference's own headers and links with its
    ColorTree and CPU engines, with nothing else from the reference's CUDA files."""
    subprocess.run(["make", "-s", "adapter"], cwd=os.path.join(ROOT, "oracle"), check=True)
    assert os.path.exists(ADAPTER)
    syms = subprocess.run(["nm", "-C", "--defined-only", ADAPTER], check=True,
                          capture_output=True, text=True).stdout
    missing = [s for s in ADAPTER_DEFINES if s not in syms]
    assert not missing, missing
    # ... and the reference's own CPU engines are the ones linked beside it
    assert "rd::RiemersmaWalkCpu(" in syms and "rd::RiemersmaBlocksCpu(" in syms
    assert "rd::ColorTree::Build(" in syms


@pytest.mark.gpu
def test_reference_adapter_matches_the_reference_cpu_engines(gpu):
    if not os.path.exists(ADAPTER):
        pytest.skip("oracle/_ref/adapter_test is not built (no reference checkout at build time)")
    r = subprocess.run([ADAPTER], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK: 0 failure(s)" in r.stdout
