"""Python access to the oracle (test infrastructure only).

Two checkers live here:

* ``liboracle.so`` -- the C restatement in ``dg_oracle.c`` of the reference's
  mesh setup, face flux, volume/lift tendency and SSP-RK3 step loop;
* ``_ref/ref_harness`` -- the reference's own Fortran modules compiled from
  source by ``oracle/Makefile`` and driven by ``ref_harness.f90``.

Only tests, the smoke check and the cpu_baseline leg of bench.py import this
module.  Arrays use the reference's Fortran (column-major) layouts; numpy
arrays are created with ``order="F"`` so that ``a.ravel(order="F")`` is the
flat layout the C and Fortran code index.
"""
from __future__ import annotations

import ctypes as C
import contextlib
import os
import shutil
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
OPERATOR_DIR = os.path.join(REPO, "operator_data")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

GENERAL = 1  # mod_dg_optr_kernel.f90:39
OPT1 = 2     # mod_dg_optr_kernel.f90:42
TYPEIDS = {"GENERAL": GENERAL, "OPT1": OPT1}

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="F_CONTIGUOUS")


def lib():
    """Load (building if needed) liboracle.so."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.path.join(HERE, "liboracle.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    L = C.CDLL(path)
    i, d, vp = C.c_int, C.c_double, C.c_void_p
    L.orc_read_operator.argtypes = [C.c_char_p, i, vp, vp, vp]
    L.orc_fmask.argtypes = [i, vp]
    L.orc_mesh_sizes.argtypes = [i, i, i, i, vp]
    L.orc_mesh_setup.argtypes = [i, i, i, i, d, d, d] + [vp] * 8
    L.orc_update_halo.argtypes = [vp, i, i, i, vp]
    L.orc_elembnd_flux.argtypes = [vp] * 9 + [i, i, i]
    L.orc_divlike.argtypes = [i, i] + [vp] * 8
    L.orc_lift.argtypes = [i, vp, vp, vp]
    L.orc_matvec_dir.argtypes = [i, i, vp, vp, vp]
    L.orc_gradlike.argtypes = [i] + [vp] * 6
    L.orc_divlike_var5.argtypes = [i] + [vp] * 8
    L.orc_cal_dqdt.argtypes = [vp] * 10 + [i, i, i]
    L.orc_cal_tend.argtypes = [vp] * 13 + [i, i, i, vp]
    L.orc_run.argtypes = [vp] * 13 + [i, i, i, i, i, d, i, vp]
    L.orc_run.restype = d
    L.orc_opt1_split.argtypes = [i]
    L.orc_opt1_split.restype = i
    L.orc_num_threads.restype = i
    _lib = L
    return L


def _p(a):
    assert a.flags.f_contiguous or a.ndim == 1
    return a.ctypes.data_as(C.c_void_p)


def read_operator(p, opdir=OPERATOR_DIR):
    """x1D(Nq), D1D(Nq,Nq), Lift_mat(Nq,Nq,Nq,6): mod_mesh.f90:253-263."""
    Nq = p + 1
    x1D = np.zeros(Nq, order="F")
    D = np.zeros((Nq, Nq), order="F")
    L = np.zeros((Nq, Nq, Nq, 6), order="F")
    rc = lib().orc_read_operator(opdir.encode(), p, _p(x1D), _p(D), _p(L))
    if rc:
        raise IOError(f"cannot read operator data p={p} in {opdir} (rc={rc})")
    return x1D, D, L


def fmask(Nq):
    F = np.zeros((Nq * Nq, 6), dtype=np.int32, order="F")
    lib().orc_fmask(Nq, _p(F))
    return F


class Mesh:
    """The reference's mesh_setup (mod_mesh.f90:74-193) on the unit cube."""

    def __init__(self, NeX, NeY, NeZ, p, dom=(1.0, 1.0, 1.0), opdir=OPERATOR_DIR):
        s = np.zeros(7, dtype=np.int32)
        lib().orc_mesh_sizes(NeX, NeY, NeZ, p, _p(s))
        (self.Nq, self.Np, self.Nfp, self.NfpTot, self.Ne, self.NeA,
         self.NhaloNode) = (int(v) for v in s)
        self.NeX, self.NeY, self.NeZ, self.p = NeX, NeY, NeZ, p
        self.x1D, self.D1D, self.Lift = read_operator(p, opdir)
        self.D1D_tr = np.asfortranarray(self.D1D.T)
        Np, Nf, Ne = self.Np, self.NfpTot, self.Ne
        self.VMapM = np.zeros((Nf, Ne), dtype=np.int32, order="F")
        self.VMapP = np.zeros((Nf, Ne), dtype=np.int32, order="F")
        self.normal_fn = np.zeros((Nf, Ne, 3), order="F")
        self.Escale = np.zeros((Np, Ne, 3), order="F")
        self.Fscale = np.zeros((Nf, Ne), order="F")
        self.pos_en = np.zeros((Np, Ne, 3), order="F")
        self.halo_src_map = np.zeros(self.NhaloNode, dtype=np.int32)
        lib().orc_mesh_setup(NeX, NeY, NeZ, p, float(dom[0]), float(dom[1]),
                             float(dom[2]), _p(self.x1D), _p(self.VMapM),
                             _p(self.VMapP), _p(self.normal_fn), _p(self.Escale),
                             _p(self.Fscale), _p(self.pos_en), _p(self.halo_src_map))

    def field(self):
        """A zeroed (Np,NeA) field."""
        return np.zeros((self.Np, self.NeA), order="F")

    def update_halo(self, f):
        lib().orc_update_halo(_p(f), self.Np, self.Ne, self.NhaloNode,
                              _p(self.halo_src_map))

    def cal_tend(self, q, u, v, w, typeid=OPT1):
        """advect3d_eq_cal_tend (mod_advect3d_eq.f90:41-85); q,u,v,w carry
        their halos.  Returns dqdt(Np,NeA) with only 1:Ne written."""
        dqdt = self.field()
        ebnd = np.zeros((self.NfpTot, self.Ne), order="F")
        lib().orc_cal_tend(_p(dqdt), _p(q), _p(u), _p(v), _p(w), _p(self.D1D),
                           _p(self.D1D_tr), _p(self.Lift), _p(self.VMapM),
                           _p(self.VMapP), _p(self.normal_fn), _p(self.Escale),
                           _p(self.Fscale), self.Nq, self.Ne, typeid, _p(ebnd))
        return dqdt

    def elembnd_flux(self, q, u, v, w):
        """cal_elembnd_flux (mod_advect3d_eq.f90:89-142)."""
        fl = np.zeros((self.NfpTot, self.Ne), order="F")
        lib().orc_elembnd_flux(_p(fl), _p(q), _p(u), _p(v), _p(w), _p(self.VMapM),
                               _p(self.VMapP), _p(self.normal_fn), _p(self.Fscale),
                               self.Np, self.NfpTot, self.Ne)
        return fl

    def run(self, q, u, v, w, dt, nstep, typeid=OPT1, minmax=False):
        """main.f90:51-80 in place on q (Np,NeA).  u,v,w must carry their
        halos.  Returns the loop wall time (and per-step min/max)."""
        mm = np.zeros((2, max(nstep, 1)), order="F") if minmax else None
        el = lib().orc_run(_p(q), _p(u), _p(v), _p(w), _p(self.D1D),
                           _p(self.D1D_tr), _p(self.Lift), _p(self.VMapM),
                           _p(self.VMapP), _p(self.normal_fn), _p(self.Escale),
                           _p(self.Fscale), _p(self.halo_src_map), self.NhaloNode,
                           self.Nq, self.Ne, self.NeA, typeid, float(dt), nstep,
                           _p(mm) if minmax else None)
        return (el, mm[:, :nstep]) if minmax else el


def divlike(Nq, typeid, D, Dtr, ix, iy, iz):
    """tensorprod_divlike_dirXYZ on one element block (Np,)."""
    ox, oy, oz = (np.zeros(Nq ** 3) for _ in range(3))
    lib().orc_divlike(Nq, typeid, _p(D), _p(Dtr), _p(ix), _p(iy), _p(iz),
                      _p(ox), _p(oy), _p(oz))
    return ox, oy, oz


def matvec_dir(Nq, direction, Mat, a):
    """dirX/dirY/dirZ (direction 0/1/2) of the OPT1 module on one block."""
    out = np.zeros(Nq ** 3)
    lib().orc_matvec_dir(Nq, direction, _p(Mat), _p(a), _p(out))
    return out


def gradlike(Nq, D, Dtr, a):
    """gradlike_dirXYZ (its intended form) on one block."""
    ox, oy, oz = (np.zeros(Nq ** 3) for _ in range(3))
    lib().orc_gradlike(Nq, _p(D), _p(Dtr), _p(a), _p(ox), _p(oy), _p(oz))
    return ox, oy, oz


def divlike_var5(Nq, D, Dtr, ix, iy, iz):
    """divlike_dirXYZ_var5 (its intended form) on one element: (Np*5,) each."""
    ox, oy, oz = (np.zeros(5 * Nq ** 3) for _ in range(3))
    lib().orc_divlike_var5(Nq, _p(D), _p(Dtr), _p(ix), _p(iy), _p(iz),
                           _p(ox), _p(oy), _p(oz))
    return ox, oy, oz


def lift(Nq, Lift, fb):
    """tensorprod_Lift_hexahedral on one element block."""
    out = np.zeros(Nq ** 3)
    lib().orc_lift(Nq, _p(Lift), _p(fb), _p(out))
    return out


# ---------------------------------------------------------------------------
# The compiled reference (oracle/_ref/ref_harness), run in a subprocess.
# ---------------------------------------------------------------------------

def ref_available():
    return os.path.exists(REF_HARNESS)


def _unlimited_stack():
    # advect3d_eq_cal_tend's automatic array ebnd_flux(NfpTot,Ne)
    # (mod_advect3d_eq.f90:66) is placed on the stack by flang: ~100 MB at
    # 32^3 elements, p=7.
    import resource
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def _ref(args, cwd, threads=None, extra_env=None):
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    env.update(extra_env or {})
    r = subprocess.run([REF_HARNESS] + [str(a) for a in args], cwd=cwd, env=env,
                       capture_output=True, text=True, preexec_fn=_unlimited_stack)
    if r.returncode:
        raise RuntimeError(f"ref_harness {args} failed: {r.stdout}\n{r.stderr}")
    return r.stdout


@contextlib.contextmanager
def _workdir():
    """Scratch directory of one reference run (its namelist, input and output
    files), removed when the run's results have been read."""
    d = tempfile.mkdtemp(prefix="dgref_")
    try:
        os.symlink(OPERATOR_DIR, os.path.join(d, "operator_data"))
        yield d
    finally:
        shutil.rmtree(d, ignore_errors=True)


def ref_mesh(NeX, NeY, NeZ, p):
    """Mesh tables as produced by the reference's mesh_setup."""
    with _workdir() as d:
        out = os.path.join(d, "mesh.bin")
        _ref(["mesh", NeX, NeY, NeZ, p, out], d)
        with open(out, "rb") as f:
            raw = f.read()
    hdr = np.frombuffer(raw[:24], dtype=np.int32)
    Nq, Np, NfpTot, Ne, NeA, nhn = (int(v) for v in hdr)
    off = 24
    res = {}

    def take(name, shape, dt):
        nonlocal off
        n = int(np.prod(shape))
        a = np.frombuffer(raw, dtype=dt, count=n, offset=off)
        off += n * np.dtype(dt).itemsize
        res[name] = a.reshape(shape, order="F").copy(order="F")

    take("x1D", (Nq,), np.float64)
    take("D1D", (Nq, Nq), np.float64)
    take("D1D_tr", (Nq, Nq), np.float64)
    take("Lift", (Nq, Nq, Nq, 6), np.float64)
    take("VMapM", (NfpTot, Ne), np.int32)
    take("VMapP", (NfpTot, Ne), np.int32)
    take("halo_src_map", (nhn,), np.int32)
    take("normal_fn", (NfpTot, Ne, 3), np.float64)
    take("Escale", (Np, Ne, 3), np.float64)
    take("Fscale", (NfpTot, Ne), np.float64)
    take("pos_en", (Np, Ne, 3), np.float64)
    res["sizes"] = (Nq, Np, NfpTot, Ne, NeA, nhn)
    return res


def ref_optr(p, tname, ix, iy, iz, fb):
    """Reference divlike + lift over nelem blocks: inputs (Np,nelem) x3 and
    (NfpTot,nelem); returns (ox, oy, oz, lift) each (Np,nelem)."""
    nelem = ix.shape[1]
    with _workdir() as d:
        inf, outf = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(inf, "wb") as f:
            for a in (ix, iy, iz, fb):
                f.write(np.asfortranarray(a, dtype=np.float64).tobytes(order="F"))
        _ref(["optr", p, tname, nelem, inf, outf], d)
        o = np.fromfile(outf, dtype=np.float64).reshape((ix.shape[0], nelem, 4), order="F")
    return tuple(np.asfortranarray(o[:, :, k]) for k in range(4))


def ref_optr2(p, a):
    """Reference dirX, dirY, dirZ and gradlike_dirXYZ (OPT1 module) over the
    blocks a(Np,nelem), with D1D for x and D1D_tr for y and z; returns
    (dx, dy, dz, gx, gy, gz) each (Np,nelem)."""
    nelem = a.shape[1]
    with _workdir() as d:
        inf, outf = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(inf, "wb") as f:
            f.write(np.asfortranarray(a, dtype=np.float64).tobytes(order="F"))
        _ref(["optr2", p, nelem, inf, outf], d)
        o = np.fromfile(outf, dtype=np.float64).reshape((a.shape[0], nelem, 6), order="F")
    return tuple(np.asfortranarray(o[:, :, k]) for k in range(6))


def write_conf(path, NeX, NeY, NeZ, p, dt, nstep, opt="OPT1", output_interval=None):
    """Writes an &PARAM_ADVECT3D namelist (main.f90:103-107) with dt as a
    17-significant-digit literal."""
    oi = output_interval or max(nstep, 1)
    with open(path, "w") as f:
        f.write("&PARAM_ADVECT3D\n")
        f.write(f"  NeX = {NeX}, NeY = {NeY}, NeZ = {NeZ},\n  PolyOrder = {p},\n")
        f.write(f'  DGOptrKernel_OptType = "{opt}"\n')
        f.write(f"  dt = {dt:.17e}".replace("e", "D") + ",\n")
        f.write(f"  nstep = {nstep},\n  output_interval = {oi},\n/\n")


def ref_bench(NeX, NeY, NeZ, p, q, u, v, w, dt, nstep, nrep, threads=None, env=None):
    """Times the reference step loop (nstep steps) nrep times from the same
    owned fields q,u,v,w (Np,Ne), with GENERAL then with OPT1, in one run of
    the harness.  Returns ({"GENERAL": [s...], "OPT1": [s...]}, threads)."""
    with _workdir() as d:
        conf, icf, outf = (os.path.join(d, n) for n in ("in.conf", "ic.bin", "out.bin"))
        write_conf(conf, NeX, NeY, NeZ, p, dt, nstep)
        with open(icf, "wb") as f:
            for a in (q, u, v, w):
                f.write(np.asfortranarray(a, dtype=np.float64).tobytes(order="F"))
        _ref(["bench", conf, icf, outf, nrep], d, threads, env)
        raw = np.fromfile(outf, dtype=np.float64)
    return {"GENERAL": raw[:nrep].tolist(), "OPT1": raw[nrep:2 * nrep].tolist()}, int(raw[-1])


def ref_run(NeX, NeY, NeZ, p, q, u, v, w, dt, nstep, opt="OPT1", mode="run",
            threads=None):
    """Runs the reference step loop (mode "run") or a single tendency
    (mode "tend") from owned fields q,u,v,w of shape (Np,Ne).  Returns
    (q_final, minmax(2,nstep), seconds) for "run" and dqdt(Np,Ne) for "tend"."""
    with _workdir() as d:
        conf, icf, outf = (os.path.join(d, n) for n in ("in.conf", "ic.bin", "out.bin"))
        write_conf(conf, NeX, NeY, NeZ, p, dt, nstep, opt)
        with open(icf, "wb") as f:
            for a in (q, u, v, w):
                f.write(np.asfortranarray(a, dtype=np.float64).tobytes(order="F"))
        stdout = _ref([mode, conf, icf, outf], d, threads)
        raw = np.fromfile(outf, dtype=np.float64)
    Np, Ne = q.shape
    qf = raw[:Np * Ne].reshape((Np, Ne), order="F")
    if mode == "tend":
        return qf
    mm = raw[Np * Ne:Np * Ne + 2 * nstep].reshape((2, nstep), order="F")
    return qf, mm, float(raw[Np * Ne + 2 * nstep]), stdout
