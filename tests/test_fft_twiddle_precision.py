"""Why the FP32 air-plane chain rounds its FFT twiddles to FP32: the twiddles are computed in FP64
(on the host) and rounded once, and with FP32 data the rounding of the data at every butterfly,
not the twiddles, sets the error.  A radix-2 FFT carried out in FP32 arithmetic on random lines
of the chain's lengths: FP64 twiddle products (the data widened, multiplied by the FP64 twiddle,
rounded back) improve the error by less than a quarter, at 1e-7, five decades below the 2e-5
plane error of the whole FP32 chain against the FP64 oracle."""
import numpy as np
import pytest


def _fft_radix2(x, tw, wide):
    n = len(x)
    bits = int(np.log2(n))
    rev = np.array([int(format(i, "0%db" % bits)[::-1], 2) for i in range(n)])
    a = x[rev].astype(np.complex64)
    s = 1
    while (1 << s) <= n:
        half, stride = 1 << (s - 1), n >> s
        pos = np.arange(half)
        w = tw[pos * stride]
        for g in range(0, n, 1 << s):
            u = a[g + pos]
            if wide:
                v = (a[g + pos + half].astype(np.complex128) * w).astype(np.complex64)
            else:
                v = a[g + pos + half] * w
            a[g + pos] = u + v
            a[g + pos + half] = u - v
        s += 1
    return a


@pytest.mark.parametrize("n", [512, 1024, 2048])
def test_fp32_twiddles_cost_little_against_fp32_data(n):
    rng = np.random.default_rng(n)
    e32, e64 = [], []
    for _ in range(3):
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        ref = np.fft.fft(x.astype(np.complex128))
        tw = np.exp(-2j * np.pi * np.arange(n) / n)
        nr = np.linalg.norm(ref)
        e32.append(np.linalg.norm(_fft_radix2(x, tw.astype(np.complex64), False) - ref) / nr)
        e64.append(np.linalg.norm(_fft_radix2(x, tw, True) - ref) / nr)
    e32, e64 = np.mean(e32), np.mean(e64)
    assert e64 < e32 < 1.25 * e64
    assert e32 < 2e-7
