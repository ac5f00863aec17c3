"""Synthetic LTI workloads (the five benchmark configurations) in numpy.

A system is drawn in FP64 and then cast to FP32, because the C ABI takes
fType = float; precision references consume the cast matrices promoted back to
double, so every path solves the same system.

Input family: u_j(t) = a sin(w t + phi) + b cos(w t + phi) + c + d t, one row of
coefficients (a, b, c, d, w, phi) per input.  The float form is registered as a
UFunc source string (ufunc_source); sampled inputs (FP64, optionally with
deterministic Gaussian noise) cover the configurations whose input is not a
pure function of float time.
"""
from dataclasses import dataclass, field

import numpy as np

RK4, ABM = 0, 1


@dataclass
class System:
    method: int
    A: np.ndarray            # n x n float32
    B: np.ndarray            # n x m float32
    C: np.ndarray            # p x n float32
    D: np.ndarray            # p x m float32
    x0: np.ndarray           # n float32
    ucoef: np.ndarray        # m x 6 float64 (a, b, c, d, w, phi)
    h: float
    steps: int               # output points
    noise_sigma: float = 0.0
    noise_seed: int = 0
    t0: float = 0.0
    tag: str = ""
    extra: dict = field(default_factory=dict)

    @property
    def n(self):
        return self.A.shape[0]

    @property
    def m(self):
        return self.B.shape[1]

    @property
    def p(self):
        return self.C.shape[0]

    @property
    def n_scan(self):
        return self.n if self.method == RK4 else 4 * self.n

    @property
    def t_end(self):
        # tEnd = tStart + (steps - 1) h, nudged by 1e-4 h so that the reference's
        # truncating step count (RungeKutta4thOrder.cpp:20) returns `steps`
        # (the SciLab stub applies the same nudge, ParODEInterfaceStubs.sci:92).
        return self.t0 + (self.steps - 1) * self.h + 1e-4 * self.h


def _orthogonal(rng, n):
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))


def make_system(method, n, m, p, lam_lo, lam_hi, h, steps, seed, skew=0.0,
                noise_sigma=0.0, tag=""):
    rng = np.random.default_rng(seed)
    Q = _orthogonal(rng, n)
    lam = rng.uniform(lam_lo, lam_hi, n)
    A = (Q * lam) @ Q.T
    if skew:
        S = rng.standard_normal((n, n)) * np.sqrt(1.0 / n)
        A = A + skew * (S - S.T) / np.sqrt(2.0)
    sd = np.sqrt(1.0 / n)
    B = rng.standard_normal((n, m)) * sd
    C = rng.standard_normal((p, n)) * sd
    D = rng.standard_normal((p, m)) * sd
    w = rng.uniform(1.0, 20.0, m)
    phi = rng.uniform(0.0, 2.0 * np.pi, m)
    x0 = rng.standard_normal(n)
    ucoef = np.zeros((m, 6))
    ucoef[:, 0] = 1.0
    ucoef[:, 4] = w.astype(np.float32)
    ucoef[:, 5] = phi.astype(np.float32)
    f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return System(method, f(A), f(B), f(C), f(D), f(x0), ucoef, h, steps,
                  noise_sigma=noise_sigma, noise_seed=seed, tag=tag)


def config(k, steps=None, seed=0):
    """Configuration k (0..2) of the benchmark set; `steps` overrides N."""
    if k == 0:
        return make_system(RK4, 8, 2, 2, -5.0, -0.1, 1e-3, steps or 4096, seed, tag="cfg0")
    if k == 1:
        return make_system(ABM, 64, 8, 8, -10.0, -0.05, 5e-4, steps or (1 << 20), seed,
                           noise_sigma=0.1, tag="cfg1")
    if k == 2:
        return make_system(RK4, 128, 16, 16, -20.0, -0.01, 2e-4, steps or (1 << 18), seed,
                           skew=0.1, tag="cfg2")
    raise ValueError(k)


def batch_config(count=512, seed=0, n_lo=4, n_hi=256, N_lo=1 << 10, N_hi=1 << 18,
                 steps_cap=None):
    """Configuration 3: `count` independent systems, RK4/ABM alternating,
    n log-uniform on [n_lo, n_hi] (integer), m = p = max(1, n // 8),
    N log-uniform on [N_lo, N_hi], h ~ U[1e-4, 1e-3]; system i from seed + i."""
    out = []
    for i in range(count):
        rng = np.random.default_rng(seed + i)
        n = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
        N = int(round(2.0 ** rng.uniform(np.log2(N_lo), np.log2(N_hi))))
        if steps_cap:
            N = min(N, steps_cap)
        h = float(rng.uniform(1e-4, 1e-3))
        m = max(1, n // 8)
        method = RK4 if i % 2 == 0 else ABM
        N = max(N, 8)
        out.append(make_system(method, n, m, m, -5.0, -0.1, h, N, seed + i + 1_000_003,
                               tag=f"cfg3[{i}]"))
    return out


def samples_f64(sysm, t0=None, count=None):
    """FP64 samples u(t0 + s h/2), s = 0 .. 2*steps (the h/2 grid), m columns."""
    t0 = sysm.t0 if t0 is None else t0
    S = 2 * sysm.steps + 1 if count is None else count
    s = np.arange(S, dtype=np.float64)
    t = t0 + s * (sysm.h / 2.0)
    a, b, c, d, w, phi = (sysm.ucoef[:, i][None, :] for i in range(6))
    g = w * t[:, None] + phi
    U = a * np.sin(g) + b * np.cos(g) + c + d * t[:, None]
    if sysm.noise_sigma:
        U = U + sysm.noise_sigma * gauss_table(sysm.noise_seed, S, sysm.m)
    return np.ascontiguousarray(U)


def gauss_table(seed, S, m):
    """N(0,1) table, row s / column j keyed by (seed, s, j): splitmix64 counter
    hashing and Box-Muller, identical to the CPU oracle's generator."""
    M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
    s = np.arange(S, dtype=np.uint64)[:, None]
    j = np.arange(m, dtype=np.uint64)[None, :]
    with np.errstate(over="ignore"):
        st = (np.uint64(seed) * np.uint64(0xD1B54A32D192ED03) + s * np.uint64(0x9E3779B97F4A7C15)
              + j * np.uint64(0xA24BAED4963EE407)) & M64

        def mix(x):
            x = x + np.uint64(0x9E3779B97F4A7C15)
            z = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return x, z ^ (z >> np.uint64(31))

        st, r1 = mix(st)
        st, r2 = mix(st)
    u1 = ((r1 >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740993.0)
    u2 = (r2 >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def ufunc_source(ucoef):
    """OpenCL-C-compatible UFunc source for the float form of the family:
    g = fma(w, t, phi); u = fma(d, t, fma(b, cos(g), fma(a, sin(g), c)))."""
    lines = ["fType UFunc(fType t, int nInputIndex){"]
    for j, (a, b, c, d, w, phi) in enumerate(np.asarray(ucoef, dtype=np.float32)):
        lit = lambda x: "%.9ef" % float(x)
        lines.append(
            f"  if(nInputIndex == {j}){{ fType g = fma((fType){lit(w)}, t, (fType){lit(phi)});"
            f" return fma((fType){lit(d)}, t, fma((fType){lit(b)}, cos(g),"
            f" fma((fType){lit(a)}, sin(g), (fType){lit(c)}))); }}")
    lines.append("  return (fType)0.0f;")
    lines.append("}")
    return "\n".join(lines) + "\n"
