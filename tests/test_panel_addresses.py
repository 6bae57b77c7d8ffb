"""CPU replay of the panel TRSM kernel's global-memory addresses
(tests/panel_replay.cpp; trsm_panel.hip with tile_stage.h).

* The stand-alone geometry of an earlier instrumented probe of this kernel
  (L and X row views of one array with ld = 8192, X at row 512, mb = 32768,
  m = 512: 1024 workgroups) — every workgroup's every access lies inside the
  array of 512 + 32768 rows, and a buffer one row short is caught.
* Every panel launch of spftrf at n = 65536 and 65535 in the four layouts,
  FP32, BF16 and split BF16 (shadow writes; panels of at most 16384 rows;
  no panel launch inside a two-leaf block except in split BF16), and at small
  ragged sizes with every workgroup replayed: reads and writes stay inside
  the RFP array, the leaf-inverse slots and the BF16 shadow, X elements
  inside the launch's mb x m block and L reads inside the strict lower
  triangle.
* The FP64 panel kernel (trsm_panel64.hip) over every panel launch of dpftrf
  at n = 32768 and 32767 in the four layouts and at small ragged sizes.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


@pytest.fixture(scope="module")
def replay(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("replay") / "panel_replay")
    cmd = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"),
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "curfp_amd", "csrc"),
           os.path.join(ROOT, "tests", "panel_replay.cpp"), "-o", exe]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0 and not os.path.isdir(os.path.join(ROCM, "include", "hip")):
        pytest.skip("HIP headers not available")
    assert res.returncode == 0, res.stderr

    def run(*args):
        return subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=120)

    return run


def test_probe_geometry_in_bounds(replay):
    r = replay("probe")
    assert r.returncode == 0, r.stdout
    # the last X element, row 512 + 32767, column 511, is the largest offset touched
    assert "array offsets [1048576, 272622079] within [0, 272629760)" in r.stdout, r.stdout


def test_probe_replay_catches_a_short_buffer(replay):
    r = replay("probe", 512 + 32768 - 1)
    assert r.returncode == 1 and "VIOLATION" in r.stdout and "X slab" in r.stdout, r.stdout


@pytest.mark.parametrize("n", [65536, 65535])
@pytest.mark.parametrize("t,u", [("N", "L"), ("N", "U"), ("T", "L"), ("T", "U")])
@pytest.mark.parametrize("bf16", [0, 1, 3])
def test_spftrf_panels_in_bounds(replay, n, t, u, bf16):
    r = replay("chain", n, t, u, bf16)
    assert r.returncode == 0, r.stdout
    assert r.stdout.startswith("ok ") and " 0 launches" not in r.stdout


@pytest.mark.parametrize("n", [129, 640, 1025, 4097, 20001])
@pytest.mark.parametrize("t,u", [("N", "U"), ("T", "L")])
@pytest.mark.parametrize("bf16", [0, 1, 3])
def test_spftrf_panels_every_workgroup(replay, n, t, u, bf16):
    r = replay("chain", n, t, u, bf16, "all")
    assert r.returncode == 0, r.stdout


@pytest.mark.parametrize("n", [32768, 32767])
@pytest.mark.parametrize("t,u", [("N", "L"), ("N", "U"), ("T", "L"), ("T", "U")])
def test_dpftrf_panels_in_bounds(replay, n, t, u):
    """FP64 panel kernel (trsm_panel64.hip): every panel launch of dpftrf."""
    r = replay("chain64", n, t, u)
    assert r.returncode == 0, r.stdout
    assert r.stdout.startswith("ok ") and "832 launches" in r.stdout, r.stdout


@pytest.mark.parametrize("n", [65, 300, 1025, 4097])
@pytest.mark.parametrize("t,u", [("N", "U"), ("T", "L")])
def test_dpftrf_panels_every_workgroup(replay, n, t, u):
    r = replay("chain64", n, t, u, "all")
    assert r.returncode == 0, r.stdout
    assert r.stdout.startswith("ok "), r.stdout
