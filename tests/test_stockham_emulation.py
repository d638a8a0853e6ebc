"""CPU emulation of the kernel FFT index maps (chromatix_amd/csrc/cxm_fft.h).

For every (N, E) pair configured in cxm_sizes.h this replays, thread by
thread and with the kernel's formulas, the register layout (v[k] <-> X[t +
k*T]), the radix schedule, the Stockham input/output/twiddle indices and the
LDS address map, and checks that the result equals numpy's FFT, that every
index stays inside its buffer and that the padded LDS map is injective.
The small DFTs (radix 2/3/4/5 and the 4x2 / 4x4 composites) are replayed
with the same decomposition and checked the same way.
"""
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
SIZES_H = ROOT / "chromatix_amd" / "csrc" / "cxm_sizes.h"



def configured_sizes():
    text = SIZES_H.read_text()
    return sorted({tuple(map(int, m)) for m in re.findall(r"X\((\d+), (\d+), (\d+), (\d+)\)", text)})


def pick_radix(rem, E):
    for r in (5, 3, 16, 8, 4, 2):
        if rem % r == 0 and E % r == 0:
            return r
    return -1


def small_dft(a, R, direction):
    """DFT_R exactly as the kernel decomposes it (composites P x Q)."""
    if R in (2, 3, 4, 5):
        W = np.exp(direction * 2j * np.pi * np.outer(np.arange(R), np.arange(R)) / R)
        return W @ a
    P, Q = {8: (4, 2), 16: (4, 4)}[R]
    b = np.empty((P, Q), complex)
    for n1 in range(P):
        b[n1] = small_dft(np.array([a[n1 + P * n2] for n2 in range(Q)]), Q, direction)
    for n1 in range(P):
        for k1 in range(Q):
            b[n1, k1] *= np.exp(direction * 2j * np.pi * n1 * k1 / R)
    out = np.empty(R, complex)
    for k1 in range(Q):
        c = small_dft(b[:, k1], P, direction)
        for k2 in range(P):
            out[k1 + Q * k2] = c[k2]
    return out


@pytest.mark.parametrize("R", [2, 3, 4, 5, 8, 16])
@pytest.mark.parametrize("direction", [-1, 1])
def test_small_dfts(R, direction):
    rng = np.random.default_rng(R)
    a = rng.standard_normal(R) + 1j * rng.standard_normal(R)
    ref = np.fft.fft(a) if direction < 0 else np.fft.ifft(a) * R
    np.testing.assert_allclose(small_dft(a, R, direction), ref, atol=1e-12)


def lds_at(e, S, N):
    """LDS slot of element e (mirror of lds_at in cxm_fft.h)."""
    if S == 1 or N & (N - 1) == 0:
        return (e + (e >> 4)) * S
    return e * S


def emulate_fft(x, N, E, direction, S=1):
    T = N // E
    L = N  # one table of length N per transform length
    tw = np.exp(-2j * np.pi * np.arange(L) / L)
    span = (N + N // 16) * S if S == 1 or N & (N - 1) == 0 else N * S
    v = np.array([[x[t + k * T] for k in range(E)] for t in range(T)], dtype=complex)
    Ns = 1
    while Ns < N:
        R = pick_radix(N // Ns, E)
        assert R > 0, (N, E)
        M = E // R
        last = Ns * R == N
        twstep = N // (Ns * R)
        lds = np.full(span, np.nan + 0j)
        written = np.zeros(span, int)
        newv = v.copy()
        for t in range(T):
            for m in range(M):
                j = t + m * T
                k = j % Ns
                a = np.array([v[t, m + r * M] for r in range(R)])
                if Ns > 1:
                    for r in range(1, R):
                        idx = k * r * twstep
                        assert idx < L
                        w = tw[idx] if direction < 0 else np.conj(tw[idx])
                        a[r] *= w
                a = small_dft(a, R, direction)
                if last:
                    for r in range(R):
                        newv[t, m + r * M] = a[r]
                else:
                    base = (j // Ns) * Ns * R + k
                    for r in range(R):
                        e = base + r * Ns
                        assert 0 <= e < N
                        addr = lds_at(e, S, N)
                        assert addr < span
                        if N & (N - 1) == 0:  # the kernels split the slot as base + constant
                            assert addr == lds_at(base, S, N) + lds_at(r * Ns, S, N)
                        lds[addr] = a[r]
                        written[addr] += 1
        if not last:
            assert written.max() == 1  # injective, no element lost
            assert written.sum() == N
            for t in range(T):
                for e in range(E):
                    newv[t, e] = lds[lds_at(t + e * T, S, N)]
                    if N & (N - 1) == 0:
                        assert lds_at(t + e * T, S, N) == lds_at(t, S, N) + lds_at(e * T, S, N)
            assert not np.isnan(newv).any()
        v = newv
        Ns *= R
    out = np.empty(N, complex)
    for t in range(T):
        for k in range(E):
            out[t + k * T] = v[t, k]
    return out


@pytest.mark.parametrize("cfg", configured_sizes(), ids=lambda c: f"N{c[0]}")
def test_configured_schedules(cfg):
    N, E_row, E_col, C = cfg
    rng = np.random.default_rng(N)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    for E, S in ((E_row, 1), (E_col, C)):
        assert N % E == 0
        # launch shape limits of the kernels
        T = N // E
        assert (T * C if S == C else max(T, 256 // max(1, T) * T)) <= 1024
        if S == C:
            span = (N + N // 16) * C if N & (N - 1) == 0 else N * C
            assert span * 8 <= 160 * 1024  # the strip fits the CU's LDS
        fwd = emulate_fft(x, N, E, -1, S)
        np.testing.assert_allclose(fwd, np.fft.fft(x), atol=1e-9 * N)
        inv = emulate_fft(x, N, E, 1, S)
        np.testing.assert_allclose(inv, np.fft.ifft(x) * N, atol=1e-9 * N)


def test_every_size_has_a_twiddle_table():
    text = (ROOT / "chromatix_amd" / "csrc" / "cxm_common.h").read_text()
    lengths = {int(v) for v in re.search(r"kTwLengths\[\] = \{([^}]*)\}", text).group(1).replace(" ", "").split(",")}
    assert {c[0] for c in configured_sizes()} <= lengths

