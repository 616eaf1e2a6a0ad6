"""CPU checks of the length-1024 three-step FFT that the FP32 film's y pass
(k_yf8) and x passes (fs8_forward / fs8_inverse) run: a lane-by-lane numpy
emulation of the kernels' index maps, twiddle powers and LDS layouts
against numpy.fft, and the layouts' bank spread.

Index maps (kernels.hip): n = b + 64 m with b = t1 + 8 t1' (m < 8 nonzero,
512 inputs), k = k2 + 128 k1 with k2 = k2' + 16 k1'; step A on lane pairs
(b, e), step B on lanes (t1 = l & 7, k2' = l >> 3), step C on lanes k2 = l.
Twiddles are powers of six per-lane bases (W = exp(-2 pi i / 1024)):
W^{8 t1' e} (W^{16 t1'})^j, W^{t1 k2'} (W^{16 t1})^j, (W^{k2})^j, (W^{8 k2'})^j.
"""
import numpy as np

N, NY = 1024, 512
W = np.exp(-2j * np.pi * np.arange(N) / N)
W16 = np.exp(-2j * np.pi * np.arange(16) / 16)


def ta(t1, t1p, k2p):
    return 4 * t1 + (k2p & 3) + 32 * (k2p >> 2) + 130 * t1p


def tb(k2, t1):
    return k2 + 132 * t1


def dft8(v, inv=False):
    j = np.arange(8)
    return np.exp((1 if inv else -1) * 2j * np.pi * np.outer(j, j) / 8) @ v


def roles(l):
    e, b = l & 1, l >> 1
    return dict(e=e, b=b, at1=b & 7, at1p=b >> 3, bt1=l & 7, bk2p=l >> 3, k2=l)


def bases(l):
    r = roles(l)
    x = [8 * r["at1p"] * r["e"], 16 * r["at1p"], r["bt1"] * r["bk2p"], 16 * r["bt1"], r["k2"], 8 * r["bk2p"]]
    return [W[v % N] for v in x]


def powers(w0, step):
    out = [w0]
    for _ in range(7):
        out.append(out[-1] * step)
    return out


def forward(x):
    ta_buf = np.zeros(1040, complex)
    tb_buf = np.zeros(1056, complex)
    for l in range(128):
        r, q = roles(l), bases(l)
        z = np.array([x[r["b"] + 64 * m] for m in range(8)])
        z = z * (W16[:8] if r["e"] else 1)
        v = dft8(z)
        wa = powers(q[0], q[1])
        for j in range(8):
            ta_buf[ta(r["at1"], r["at1p"], 2 * j + r["e"])] = v[j] * wa[j]
    for l in range(128):
        r, q = roles(l), bases(l)
        v = dft8(np.array([ta_buf[ta(r["bt1"], j, r["bk2p"])] for j in range(8)]))
        wb = powers(q[2], q[3])
        for j in range(8):
            tb_buf[tb(r["bk2p"] + 16 * j, r["bt1"])] = v[j] * wb[j]
    out = np.zeros(N, complex)
    for k2 in range(128):
        v = dft8(np.array([tb_buf[tb(k2, j)] for j in range(8)]))
        for j in range(8):
            out[k2 + 128 * j] = v[j]
    return out


def inverse(X):
    ta_buf = np.zeros(1040, complex)
    tb_buf = np.zeros(1056, complex)
    for l in range(128):
        r, q = roles(l), bases(l)
        v = dft8(np.array([X[r["k2"] + 128 * j] for j in range(8)]), inv=True)
        wc = powers(1.0, q[4])
        for j in range(8):
            tb_buf[tb(r["k2"], j)] = v[j] * np.conj(wc[j])
    for l in range(128):
        r, q = roles(l), bases(l)
        v = dft8(np.array([tb_buf[tb(r["bk2p"] + 16 * j, r["bt1"])] for j in range(8)]), inv=True)
        wd = powers(1.0, q[5])
        for j in range(8):
            ta_buf[ta(r["bt1"], j, r["bk2p"])] = v[j] * np.conj(wd[j])
    half = {}
    for l in range(128):
        r = roles(l)
        v = dft8(np.array([ta_buf[ta(r["at1"], r["at1p"], 2 * j + r["e"])] for j in range(8)]), inv=True)
        half[l] = v * (np.conj(W16[:8]) if r["e"] else 1)
    out = np.zeros(NY, complex)
    for l in range(128):
        r = roles(l)
        for i in range(4):
            m = 2 * i + r["e"]
            out[r["b"] + 64 * m] = half[l][m] + half[l ^ 1][m]
    return out


def test_forward_and_inverse_equal_numpy_fft():
    rng = np.random.default_rng(11)
    x = rng.standard_normal(NY) + 1j * rng.standard_normal(NY)
    X = forward(x)
    ref = np.fft.fft(np.concatenate([x, np.zeros(N - NY)]))
    assert np.abs(X - ref).max() <= 1e-12 * np.abs(ref).max()
    Y = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    y = inverse(Y)
    ref = np.fft.ifft(Y)[:NY] * N
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()


def test_layouts_fit_and_are_injective():
    a = [ta(t1, t1p, k2p) for t1 in range(8) for t1p in range(8) for k2p in range(16)]
    b = [tb(k2, t1) for k2 in range(128) for t1 in range(8)]
    assert len(set(a)) == len(a) and max(a) < 1040
    assert len(set(b)) == len(b) and max(b) < 1056


def test_layouts_spread_each_half_wave_over_all_bank_pairs():
    # complex FP32 values span two 4-byte banks; a half-wave's 32 lanes hit
    # 32 distinct bank pairs when their indices differ mod 32
    def distinct(idx):
        return len({i % 32 for i in idx}) == len(idx)
    for hw in range(4):
        lanes = range(32 * hw, 32 * hw + 32)
        for j in range(8):
            assert distinct([ta((l >> 1) & 7, (l >> 1) >> 3, 2 * j + (l & 1)) for l in lanes])   # A write, A' read
            assert distinct([ta(l & 7, j, l >> 3) for l in lanes])                              # B read, B' write
            assert distinct([tb((l >> 3) + 16 * j, l & 7) for l in lanes])                      # B write, B' read
            assert distinct([tb(l, j) for l in lanes])                                          # C read, C' write
