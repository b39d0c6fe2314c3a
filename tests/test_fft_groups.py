"""The lane-group FFTs of the 2-D/3-D kernels (csrc/mqpiv_fft.hpp), run on the CPU
by tests/fft_groups_emu.cpp (every lane of a group emulated, exchanges replayed)
and compared with a direct DFT: forward (decimation in time, strided in, chunked
out) and inverse (decimation in frequency, chunked in, strided out) for every
(length, lanes per line) the kernels instantiate."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def _run(tmp_path_factory, *defines):
    if not Path(HIPCC).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path_factory.mktemp("fftemu") / "fft_groups_emu"
    # host code only: the header's device paths are not compiled for the CPU run
    r = subprocess.run([HIPCC, "-O1", "-std=c++17", *defines, str(HERE / "fft_groups_emu.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True, timeout=300).stdout
    res = {}
    for line in out.strip().splitlines():
        L, P, tn, d, err = line.split()
        res[(int(L), int(P), tn, d)] = float(err)
    return res


@pytest.fixture(scope="module")
def emu_results(tmp_path_factory):
    return _run(tmp_path_factory)


def test_every_group_matches_the_dft(emu_results):
    assert len(emu_results) == 58
    for (L, P, tn, d), err in emu_results.items():
        # inputs U(-1, 1); rounding grows slowly with L
        tol = 5e-6 if tn == "f32" else 5e-14
        assert err < tol * (L / 64) ** 0.5 * 4, (L, P, tn, d, err)


def test_a_wrong_partner_is_detected(tmp_path_factory):
    """Planted mistake: every exchange reads the lane next to its partner.  The
    same comparison must then fail in every case (errors O(1), not rounding)."""
    bad = _run(tmp_path_factory, "-DEMU_WRONG_PARTNER")
    assert len(bad) == 58
    assert min(bad.values()) > 1e-2, min(bad.items(), key=lambda kv: kv[1])
