"""Checkpoint files in the reference's format (W/checkpoint.f90:12-55,
W/particle_initialization.f90:22-82, names from W/basic_functions.fh:52-87).

Per image k (here always 1) and redshift z the state is six raw stream
files under opath/image<k>/:  <z>_info_<k>.bin (the sim_header record),
<z>_xp_<k>.bin, <z>_vp_<k>.bin, <z>_np_<k>.bin (rhoc of the physical cells),
<z>_vc_<k>.bin (vfield), <z>_id_<k>.bin (pid), with <z> written as
Fortran's (f7.3) and left-adjusted.  Arrays are Fortran column-major, which
is the cell/particle order the library uses.
"""
import os

import numpy as np

# sim_header (W/parameters.f90:136-160): 17 integer(8) then 22 real(4)
INT_FIELDS = ["nplocal", "npglobal", "nplocal_nu", "npglobal_nu", "izipx", "izipv", "izipx_nu",
              "izipv_nu", "image", "nn", "nnt", "nt", "ncell", "ncb", "timestep", "cur_checkpoint",
              "cur_halofind"]
REAL_FIELDS = ["a", "t", "tau", "dt_pp", "dt_fine", "dt_coarse", "dt_vmax", "dt_vmax_nu", "mass_p_cdm",
               "mass_p_nu", "box", "h0", "omega_m", "omega_l", "s8", "vsim2phys", "sigma_vres", "sigma_vi",
               "sigma_vi_nu", "z_i", "z_i_nu", "vz_max"]
HEADER_DTYPE = np.dtype([(n, "<i8") for n in INT_FIELDS] + [(n, "<f4") for n in REAL_FIELDS])
assert HEADER_DTYPE.itemsize == 224


def zstr(z):
    """Fortran write(str,'(f7.3)') z; trim(adjustl(str))."""
    s = "%7.3f" % float(np.float32(z))
    if len(s) > 7:
        s = "*" * 7
    return s.strip()


def names(opath, z, image=1):
    d = os.path.join(opath, f"image{image}")
    pre = os.path.join(d, f"{zstr(z)}_")
    suf = f"_{image}.bin"
    return {k: pre + k + suf for k in ("info", "xp", "vp", "np", "vc", "id")}


def new_header(**kw):
    h = np.zeros((), HEADER_DTYPE)
    for k, v in kw.items():
        h[k] = v
    return h


def write(opath, z, header, xp, vp, rhoc, vfield, pid=None, image=1):
    """Write one checkpoint (W/checkpoint.f90:370-400)."""
    fn = names(opath, z, image)
    os.makedirs(os.path.dirname(fn["info"]), exist_ok=True)
    h = np.array(header, HEADER_DTYPE)
    h["nplocal"] = xp.shape[0]
    h.tofile(fn["info"])
    # the files hold izipx/izipv bytes per component, whatever the caller's dtype
    xt = np.int8 if h["izipx"] == 1 else np.int16
    vt = np.int8 if h["izipv"] == 1 else np.int16
    for name, a, t in (("xp", xp, xt), ("vp", vp, vt)):
        a = np.asarray(a)
        if a.size and (a.min() < np.iinfo(t).min or a.max() > np.iinfo(t).max):
            raise ValueError(f"{name}: values outside {np.dtype(t).name}")
        np.ascontiguousarray(a, t).tofile(fn[name])
    np.ascontiguousarray(rhoc, np.int32).tofile(fn["np"])
    np.ascontiguousarray(vfield, np.float32).tofile(fn["vc"])
    if pid is not None:
        np.ascontiguousarray(pid, np.int32).tofile(fn["id"])
    return fn


def read(opath, z, nc, image=1):
    """Read one checkpoint (W/particle_initialization.f90:479-509)."""
    fn = names(opath, z, image)
    h = np.fromfile(fn["info"], HEADER_DTYPE, count=1)[0]
    n = int(h["nplocal"])
    xt = np.int8 if h["izipx"] == 1 else np.int16
    vt = np.int8 if h["izipv"] == 1 else np.int16
    xp = np.fromfile(fn["xp"], xt, count=3 * n).reshape(n, 3)
    vp = np.fromfile(fn["vp"], vt, count=3 * n).reshape(n, 3)
    rhoc = np.fromfile(fn["np"], np.int32, count=nc ** 3)
    vf = np.fromfile(fn["vc"], np.float32, count=3 * nc ** 3).reshape(-1, 3)
    pid = np.fromfile(fn["id"], np.int32, count=n) if os.path.exists(fn["id"]) else None
    return h, dict(xp=xp, vp=vp, rhoc=rhoc, vfield=vf, pid=pid, nplocal=n)
