"""ctypes binding of the differentiable PM (include/cube_dpm.h) and its host
schedule: growth, ICs and KDK coefficients in FP64.

Units: positions in cells of the nf^3 mesh, time variable a, H0 = 1.
  dx/da = p / (a^3 E(a)),   dp/da = (3/2) Om F(x) / (a^2 E(a)),
  F = -grad phi with lap phi = delta (mesh units).
Zel'dovich ICs at a0: x = q + D(a0) psi, p = a0^2 D(a0) f(a0) E(a0) psi, with
psi the linear displacement at D=1 (lattice units, scaled to mesh cells by
the library).  The KDK step a_s -> a_{s+1} uses
  kA = 1.5 Om int_{a_s}^{a_mid} da/(a^2 E), d = int_{a_s}^{a_{s+1}} da/(a^3 E),
  kB = 1.5 Om int_{a_mid}^{a_{s+1}} da/(a^2 E).
"""
import ctypes

import numpy as np
from scipy.integrate import quad

from . import cube, ics


class DPMParams(ctypes.Structure):
    _fields_ = [("ng", ctypes.c_int32), ("nf", ctypes.c_int32), ("nsteps", ctypes.c_int32),
                ("checkpoint", ctypes.c_int32), ("device", ctypes.c_int32), ("reserved", ctypes.c_int32)]


EXPORTS = ["cube_dpm_create", "cube_dpm_destroy", "cube_dpm_set_amplitude", "cube_dpm_set_schedule",
           "cube_dpm_forward", "cube_dpm_adjoint", "cube_dpm_state", "cube_dpm_sync"]


def lib():
    L = cube.lib()
    if not getattr(L, "_dpm_bound", False):
        vp = ctypes.c_void_p
        L.cube_dpm_create.argtypes = [ctypes.POINTER(DPMParams), ctypes.POINTER(vp)]
        L.cube_dpm_destroy.argtypes = [vp]
        L.cube_dpm_destroy.restype = None
        L.cube_dpm_set_amplitude.argtypes = [vp, vp]
        L.cube_dpm_set_schedule.argtypes = [vp, vp, vp]
        L.cube_dpm_forward.argtypes = [vp, vp, vp]
        L.cube_dpm_adjoint.argtypes = [vp, vp, vp]
        L.cube_dpm_state.argtypes = [vp, vp, vp]
        L.cube_dpm_sync.argtypes = [vp]
        L._dpm_bound = True
    return L


def _hptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ----------------------------------------------------------------- schedule
def efunc(a, om):
    return np.sqrt(om / a ** 3 + (1.0 - om))


def growth_rate(a, om):
    """f = dln D/dln a of the growing mode (numerical derivative of ics.dgrow)."""
    h = 1e-5 * a
    return (np.log(ics.dgrow(a + h, om, 1 - om)) - np.log(ics.dgrow(a - h, om, 1 - om))) / \
        (np.log(a + h) - np.log(a - h))


def schedule(a0, a1, nsteps, om=0.31):
    """ic=(cx, cp) and steps[nsteps,3]=(kA, d, kB), linear in a; growth
    normalised to D(a=1)=1."""
    D1 = ics.dgrow(1.0, om, 1 - om)
    D = ics.dgrow(a0, om, 1 - om) / D1
    f = growth_rate(a0, om)
    ic = np.array([D, a0 ** 2 * D * f * efunc(a0, om)])
    aa = np.linspace(a0, a1, nsteps + 1)
    st = np.zeros((nsteps, 3))
    for s in range(nsteps):
        am = 0.5 * (aa[s] + aa[s + 1])
        st[s, 0] = 1.5 * om * quad(lambda a: 1 / (a * a * efunc(a, om)), aa[s], am, epsabs=0, epsrel=1e-12)[0]
        st[s, 1] = quad(lambda a: 1 / (a ** 3 * efunc(a, om)), aa[s], aa[s + 1], epsabs=0, epsrel=1e-12)[0]
        st[s, 2] = 1.5 * om * quad(lambda a: 1 / (a * a * efunc(a, om)), am, aa[s + 1], epsabs=0, epsrel=1e-12)[0]
    return ic, st


def amplitude(ng, box, cosmo=None):
    """A(k) = sqrt(P(k)/V_cell) of the linear density (z=0) on the lattice half
    spectrum [kz][ky][kx], so delta_k = A w_k for unit white noise w."""
    cosmo = cosmo or ics.Cosmology()
    m = np.fft.fftfreq(ng) * ng
    mx = np.arange(ng // 2 + 1)
    KZ, KY, KX = np.meshgrid(m, m, mx, indexing="ij")
    k = 2 * np.pi / box * np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    pk = np.where(k > 0, cosmo.pk(np.where(k > 0, k, 1.0)), 0.0)
    return np.sqrt(pk / (box / ng) ** 3).astype(np.float32)


# ------------------------------------------------------------------ handle
class DPM:
    """One box of the differentiable PM on the GPU.  forward/adjoint take
    device pointers (ints) of float32 arrays."""

    def __init__(self, ng, nf, nsteps, checkpoint=True, device=0):
        self.ng, self.nf, self.nsteps = ng, nf, nsteps
        self._h = ctypes.c_void_p()
        p = DPMParams(ng, nf, nsteps, int(checkpoint), device, 0)
        self._check(lib().cube_dpm_create(ctypes.byref(p), ctypes.byref(self._h)), "cube_dpm_create")

    def close(self):
        if self._h:
            lib().cube_dpm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _check(st, what):
        if st:
            raise cube.CubeError(st, what, lib().cube_strerror(st).decode())

    def set_amplitude(self, amp):
        a = np.ascontiguousarray(amp, np.float32)
        assert a.size == self.ng * self.ng * (self.ng // 2 + 1)
        self._check(lib().cube_dpm_set_amplitude(self._h, _hptr(a)), "set_amplitude")

    def set_schedule(self, ic, steps):
        ic = np.ascontiguousarray(ic, np.float64)
        steps = np.ascontiguousarray(steps, np.float64)
        assert steps.shape == (self.nsteps, 3)
        self._check(lib().cube_dpm_set_schedule(self._h, _hptr(ic), _hptr(steps)), "set_schedule")

    def forward(self, noise_ptr, delta_ptr):
        self._check(lib().cube_dpm_forward(self._h, noise_ptr, delta_ptr), "forward")

    def adjoint(self, delta_bar_ptr, noise_bar_ptr):
        self._check(lib().cube_dpm_adjoint(self._h, delta_bar_ptr, noise_bar_ptr), "adjoint")

    def state(self, x_ptr, p_ptr):
        self._check(lib().cube_dpm_state(self._h, x_ptr, p_ptr), "state")

    def sync(self):
        self._check(lib().cube_dpm_sync(self._h), "sync")


# ------------------------------------------------------------ ensemble runner
BATCH_EXPORTS = ["cube_batch_create", "cube_batch_destroy", "cube_batch_box", "cube_batch_set_io",
                 "cube_batch_tune", "cube_batch_run", "cube_batch_sync"]


def _bind_batch(L):
    if getattr(L, "_batch_bound", False):
        return L
    vp = ctypes.c_void_p
    L.cube_batch_create.argtypes = [ctypes.c_int32, ctypes.POINTER(DPMParams), ctypes.POINTER(vp)]
    L.cube_batch_destroy.argtypes = [vp]
    L.cube_batch_destroy.restype = None
    L.cube_batch_box.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(vp)]
    L.cube_batch_set_io.argtypes = [vp, ctypes.c_int32, vp, vp, vp, vp, vp]
    L.cube_batch_tune.argtypes = [vp, vp]
    L.cube_batch_run.argtypes = [vp, ctypes.c_int32]
    L.cube_batch_sync.argtypes = [vp]
    L._batch_bound = True
    return L


class _BoxView(DPM):
    """A box owned by a Batch (not destroyed on its own)."""

    def __init__(self, handle, ng, nf, nsteps):
        self.ng, self.nf, self.nsteps = ng, nf, nsteps
        self._h = handle

    def close(self):
        self._h = ctypes.c_void_p()


class Batch:
    """Ensemble of differentiable-PM boxes (include/cube_batch.h)."""

    def __init__(self, specs, checkpoint=True, device=0):
        """specs: list of (ng, nf, nsteps)."""
        L = _bind_batch(lib())
        arr = (DPMParams * len(specs))(*[DPMParams(ng, nf, ns, int(checkpoint), device, 0) for ng, nf, ns in specs])
        self._h = ctypes.c_void_p()
        DPM._check(L.cube_batch_create(len(specs), arr, ctypes.byref(self._h)), "cube_batch_create")
        self.specs = list(specs)
        self.boxes = []
        for i, (ng, nf, ns) in enumerate(specs):
            h = ctypes.c_void_p()
            DPM._check(L.cube_batch_box(self._h, i, ctypes.byref(h)), "cube_batch_box")
            self.boxes.append(_BoxView(h, ng, nf, ns))

    def set_io(self, i, noise, target, delta, noise_bar, loss):
        DPM._check(_bind_batch(lib()).cube_batch_set_io(self._h, i, noise, target, delta, noise_bar, loss),
                   "cube_batch_set_io")

    def tune(self):
        shapes = np.zeros(2 * len(self.specs), np.int32)
        DPM._check(_bind_batch(lib()).cube_batch_tune(self._h, _hptr(shapes)), "cube_batch_tune")
        return shapes.reshape(-1, 2)

    def run(self, graph=True):
        DPM._check(_bind_batch(lib()).cube_batch_run(self._h, 1 if graph else 0), "cube_batch_run")

    def sync(self):
        DPM._check(_bind_batch(lib()).cube_batch_sync(self._h), "cube_batch_sync")

    def close(self):
        if self._h:
            for b in self.boxes:
                b.close()
            _bind_batch(lib()).cube_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_ensemble(classes=((32, 16), (64, 16), (96, 8), (128, 8)), nsteps=10, seed0=1000, a0=0.1, a1=1.0,
                  checkpoint=True):
    """The uneven ensemble of boxes: `classes` = (particles per dim, count);
    fine mesh 2x the particle grid, each box with its own seed (seed0 + i),
    box size U(50, 200) Mpc/h and sigma8 U(0.7, 0.9) drawn from that seed,
    white-noise ICs and an MSE target density from a different seed.
    Returns (batch, tensors) with the device tensors wired to the batch."""
    import torch
    specs, meta = [], []
    i = 0
    for ng, count in classes:
        for _ in range(count):
            rng = np.random.default_rng(seed0 + i)
            meta.append(dict(seed=seed0 + i, box=float(rng.uniform(50, 200)), sigma8=float(rng.uniform(0.7, 0.9))))
            specs.append((ng, 2 * ng, nsteps))
            i += 1
    b = Batch(specs, checkpoint=checkpoint)
    ic, steps = schedule(a0, a1, nsteps)
    t = []
    for (ng, nf, ns), m, box in zip(specs, meta, b.boxes):
        box.set_amplitude(amplitude(ng, m["box"], ics.Cosmology(sigma8=m["sigma8"])))
        box.set_schedule(ic, steps)
        rng = np.random.default_rng(m["seed"])
        noise = torch.from_numpy(rng.standard_normal(ng ** 3).astype(np.float32)).cuda()
        target = torch.from_numpy((np.random.default_rng(m["seed"] + 7919).standard_normal(nf ** 3) * 0.5)
                                  .astype(np.float32)).cuda()
        tt = dict(noise=noise, target=target, delta=torch.zeros(nf ** 3, device="cuda"),
                  noise_bar=torch.zeros(ng ** 3, device="cuda"),
                  loss=torch.zeros(1, dtype=torch.float64, device="cuda"), **m)
        t.append(tt)
    torch.cuda.synchronize()
    for k, tt in enumerate(t):
        b.set_io(k, tt["noise"].data_ptr(), tt["target"].data_ptr(), tt["delta"].data_ptr(),
                 tt["noise_bar"].data_ptr(), tt["loss"].data_ptr())
    return b, t
