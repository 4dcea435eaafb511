"""The per-body FFT's plans and index arithmetic, checked on the CPU.

tests/fft_emul.cpp runs the stage functions of electricbubbles_amd/csrc/eb_fft.h with the
plans and tables of eb_fft_plan.h in a sequential emulation of one workgroup (all loads of a
stage, then all stores).  Checked against numpy's FFT, which has DCFFTF's definition
(dcfft.f:45-51): every length 2..3200 forward, a sample backward, the Bluestein and the
direct-sum paths for the same lengths, and the magic-number division over its whole domain
of use.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "electricbubbles_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("fftemu") / "fft_emul.so")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", "-I", CSRC, os.path.join(HERE, "fft_emul.cpp"),
                    "-o", out], check=True, capture_output=True)
    lib = ctypes.CDLL(out)
    lib.fft_emul.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                             ctypes.c_void_p]
    lib.fdiv_emul.restype = ctypes.c_int
    lib.fdiv_emul.argtypes = [ctypes.c_int, ctypes.c_int]
    return lib


def transform(lib, lens, sign, budget=5888, seed=1):
    lens = np.asarray(lens, np.int32)
    rng = np.random.default_rng(seed)
    data = [rng.normal(size=n) + 1j * rng.normal(size=n) for n in lens]
    buf = np.concatenate(data).astype(np.complex128)
    info = np.zeros(4, np.int32)
    lib.fft_emul(len(lens), lens.ctypes.data, buf.ctypes.data, sign, budget, info.ctypes.data)
    off = np.concatenate([[0], np.cumsum(lens)])
    errs = []
    for k, (n, d) in enumerate(zip(lens, data)):
        ref = np.fft.fft(d) if sign < 0 else n * np.fft.ifft(d)
        errs.append(np.linalg.norm(buf[off[k]:off[k + 1]] - ref) / np.linalg.norm(ref))
    return np.array(errs), info


def test_every_length_forward(emu):
    errs, _ = transform(emu, np.arange(2, 3201), -1)
    assert errs.max() < 1e-14, int(np.argmax(errs)) + 2


def test_backward_sample(emu):
    lens = np.concatenate([np.arange(2, 300), np.random.default_rng(3).integers(300, 3201, 300)])
    errs, _ = transform(emu, lens, +1)
    assert errs.max() < 1e-14


@pytest.mark.parametrize("n,P,Ns,M", [(1942, 971, 2, 1944), (2026, 1013, 2, 2025), (578, 17, 34, 36),
                                      (1009, 1009, 1, 2025), (17, 17, 1, 36)])
def test_bluestein_plans(emu, n, P, Ns, M):
    """The largest prime factor of at least 17 becomes a Bluestein stage with the smallest
    2^a 3^b 5^c length M >= 2P - 1; with no workspace it stays a direct sum."""
    errs, info = transform(emu, [n], -1)
    assert tuple(info[:3]) == (P, Ns, M) and errs[0] < 1e-14
    errs0, info0 = transform(emu, [n], -1, budget=0)
    assert info0[0] == 0 and errs0[0] < 1e-14


def test_closed_form_radices_exact_on_unit_impulses(emu):
    """DFT of e_j is the column exp(-2 pi i j k / n): checks every stage's index map."""
    for n in (8, 12, 30, 40, 64, 96, 120, 2048):
        for j in (0, 1, n // 3, n - 1):
            lens = np.array([n], np.int32)
            buf = np.zeros(n, np.complex128)
            buf[j] = 1.0
            info = np.zeros(4, np.int32)
            emu.fft_emul(1, lens.ctypes.data, buf.ctypes.data, -1, 5888, info.ctypes.data)
            ref = np.exp(-2j * np.pi * ((j * np.arange(n)) % n) / n)  # exact argument reduction
            assert np.abs(buf - ref).max() < 1e-15 * 8, (n, j)


def test_magic_division(emu):
    """t / d by umulhi with m = floor(2^32/d) + 1 is exact for t d < 2^32; used for t, d <= 8192."""
    rng = np.random.default_rng(4)
    for d in list(range(1, 65)) + list(rng.integers(65, 8193, 200)) + [4096, 8191, 8192]:
        d = int(d)
        ts = np.unique(np.concatenate([np.arange(0, 8193, max(1, d // 7)), np.arange(d - 1, 8193, d),
                                       np.arange(d, 8193, d), rng.integers(0, 8193, 200)]))
        for t in ts:
            assert emu.fdiv_emul(int(t), d) == int(t) // d, (t, d)
