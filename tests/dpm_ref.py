"""FP64 CPU twin of the differentiable PM (include/cube_dpm.h), forward and
hand-derived adjoint, for checking the GPU kernels.  Test infrastructure:
the same operators as the library in float64 numpy, written independently
of the HIP code.  Conventions match the library: positions in mesh cells,
CIC on points floor(x), floor(x)+1; K_d = k_d/k^2, zero at k=0 and on the
Nyquist plane of axis d."""
import numpy as np


def kfac(n, d, half=True):
    m = np.fft.fftfreq(n) * n
    mx = np.arange(n // 2 + 1) if half else m
    MZ, MY, MX = np.meshgrid(m, m, mx, indexing="ij")
    M = [MX, MY, MZ]
    k = [2 * np.pi / n * a for a in M]
    k2 = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
    with np.errstate(invalid="ignore", divide="ignore"):
        K = np.where(k2 > 0, k[d] / np.where(k2 > 0, k2, 1), 0.0)
    K = np.where(np.abs(M[d]) * 2 == n, 0.0, K)
    return K


class Ref:
    def __init__(self, ng, nf, amp, ic, steps):
        self.ng, self.nf = ng, nf
        self.amp = np.asarray(amp, np.float64)
        self.ic = np.asarray(ic, np.float64)
        self.steps = np.asarray(steps, np.float64)
        self.Kg = [kfac(ng, d) for d in range(3)]
        self.Kf = [kfac(nf, d) for d in range(3)]
        r = nf / ng
        l = np.arange(ng)
        LZ, LY, LX = np.meshgrid(l, l, l, indexing="ij")
        self.q = np.stack([LX.ravel(), LY.ravel(), LZ.ravel()], axis=1) * r

    # -- CIC ----------------------------------------------------------------
    def _cic(self, x):
        n = self.nf
        fl = np.floor(x)
        f = x - fl
        i0 = fl.astype(np.int64) % n
        i1 = (i0 + 1) % n
        return (i0, i1), (1 - f, f)

    def _corners(self, x):
        (i0, i1), (w0, w1) = self._cic(x)
        n = self.nf
        for cz in (0, 1):
            for cy in (0, 1):
                for cx in (0, 1):
                    ix = i1[:, 0] if cx else i0[:, 0]
                    iy = i1[:, 1] if cy else i0[:, 1]
                    iz = i1[:, 2] if cz else i0[:, 2]
                    w = [w1[:, 0] if cx else w0[:, 0], w1[:, 1] if cy else w0[:, 1], w1[:, 2] if cz else w0[:, 2]]
                    s = [1.0 if cx else -1.0, 1.0 if cy else -1.0, 1.0 if cz else -1.0]
                    yield (iz * n + iy) * n + ix, w, s

    def deposit(self, x, val=None):
        n = self.nf
        rho = np.zeros(n ** 3)
        for m, w, _ in self._corners(x):
            ww = w[0] * w[1] * w[2]
            rho += np.bincount(m, weights=ww if val is None else ww * val, minlength=n ** 3)
        return rho.reshape(n, n, n)

    def gather3(self, x, F):
        out = np.zeros((x.shape[0], 3))
        Ff = F.reshape(-1, 3)
        for m, w, _ in self._corners(x):
            out += (w[0] * w[1] * w[2])[:, None] * Ff[m]
        return out

    def grad(self, x, lam=None, s_lam=1.0, g=None, M=None):
        """sum_c dW/dx (s_lam lam_c + g . M_c)"""
        out = np.zeros((x.shape[0], 3))
        lf = None if lam is None else lam.ravel()
        Mf = None if M is None else M.reshape(-1, 3)
        for m, w, s in self._corners(x):
            v = np.zeros(x.shape[0])
            if lf is not None:
                v += s_lam * lf[m]
            if Mf is not None:
                v += (g * Mf[m]).sum(1)
            out[:, 0] += v * s[0] * w[1] * w[2]
            out[:, 1] += v * w[0] * s[1] * w[2]
            out[:, 2] += v * w[0] * w[1] * s[2]
        return out

    # -- operators ----------------------------------------------------------
    def force(self, x):
        n = self.nf
        rk = np.fft.rfftn(self.deposit(x))
        F = np.stack([np.fft.irfftn(1j * self.Kf[d] * rk, s=(n, n, n), axes=(0, 1, 2)) for d in range(3)], axis=-1)
        return F  # [z][y][x][d]

    def ics(self, w):
        ng, r = self.ng, self.nf / self.ng
        wk = np.fft.rfftn(w.reshape(ng, ng, ng))
        psi = np.stack([np.fft.irfftn(1j * self.Kg[d] * self.amp * wk, s=(ng, ng, ng), axes=(0, 1, 2)).ravel() for d in range(3)],
                       axis=1)
        x = np.mod(self.q + self.ic[0] * psi * r, self.nf)
        p = self.ic[1] * psi * r
        return x, p

    def forward(self, w):
        x, p = self.ics(w)
        xs = [x.copy()]
        F = self.gather3(x, self.force(x))
        for kA, d, kB in self.steps:
            p = p + kA * F
            x = np.mod(x + d * p, self.nf)
            xs.append(x.copy())
            F = self.gather3(x, self.force(x))
            p = p + kB * F
        nbar = x.shape[0] / self.nf ** 3
        delta = self.deposit(x) / nbar - 1.0
        self.xs = xs
        return delta

    def vjp_force(self, x, g):
        """d/dx of sum_i g_i . F_i(x)"""
        n = self.nf
        M = self.force(x)
        lam_k = 0
        for d in range(3):
            Sk = np.fft.rfftn(self.deposit(x, g[:, d]))
            lam_k = lam_k + (-1j) * self.Kf[d] * Sk
        lam = np.fft.irfftn(lam_k, s=(n, n, n), axes=(0, 1, 2))
        return self.grad(x, lam=lam, s_lam=1.0, g=g, M=M)

    def adjoint(self, dbar):
        ng, nf, r = self.ng, self.nf, self.nf / self.ng
        xs = self.xs
        nbar = xs[-1].shape[0] / nf ** 3
        xbar = self.grad(xs[-1], lam=dbar, s_lam=1.0 / nbar)
        pbar = np.zeros_like(xbar)
        S = len(self.steps)
        for s in range(S - 1, -1, -1):
            pbar = pbar + self.steps[s, 1] * xbar
            coef = self.steps[s, 0] + (self.steps[s - 1, 2] if s > 0 else 0.0)
            xbar = xbar + self.vjp_force(xs[s], coef * pbar)
        psib = r * (self.ic[0] * xbar + self.ic[1] * pbar)
        acc = 0
        for d in range(3):
            acc = acc + (-1j) * self.Kg[d] * self.amp * np.fft.rfftn(psib[:, d].reshape(ng, ng, ng))
        return np.fft.irfftn(acc, s=(ng, ng, ng), axes=(0, 1, 2)).ravel()
