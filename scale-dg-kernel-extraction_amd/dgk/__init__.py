"""ctypes binding of libdgk.so (include/dgk.h) for Python hosts and tests.

Mirrors the reference's operator interface: ``optr_kernel_setup`` is
dg_optr_kernel_setup (mod_dg_optr_kernel.f90:45-58), ``mesh_setup`` is the
reference's mesh_setup (mod_mesh.f90:74-193), ``Solver.cal_tend`` is
advect3d_eq_cal_tend (mod_advect3d_eq.f90:41-85) and ``Solver.step`` one
SSP-RK3 step of main.f90:51-80.  Arrays follow the reference's Fortran
layouts (numpy ``order="F"``).  The library is required: nothing here falls
back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "lib", "libdgk.so")
OPERATOR_DIR = os.path.join(REPO, "operator_data")

OK = 0
ERR_ARG, ERR_OPTTYPE, ERR_ORDER, ERR_IO, ERR_HIP, ERR_NOMEM, ERR_STATE, ERR_NONFINITE = range(1, 9)
TYPE_GENERAL, TYPE_OPT1, TYPE_HIP = 1, 2, 3
FLAG_FMA, FLAG_FORCE_TABLES, FLAG_NODE_KERNEL, FLAG_FACTORED_LIFT = 0x1, 0x2, 0x4, 0x8
FLAG_CONST_VELOCITY = 0x10
FLAG_MIXED = 0x20
FLAG_MFMA = 0x40
FIELD_Q, FIELD_Q0, FIELD_DQDT, FIELD_U, FIELD_V, FIELD_W = range(6)

EXPORTS = (
    "dgk_last_error", "dgk_version", "dgk_optr_kernel_setup", "dgk_read_operator",
    "dgk_mesh_sizes", "dgk_mesh_setup", "dgk_create", "dgk_create_structured",
    "dgk_destroy", "dgk_set_fields", "dgk_get_field", "dgk_field_ptr", "dgk_cal_tend",
    "dgk_rk_stage", "dgk_step", "dgk_run", "dgk_synchronize", "dgk_query",
    "dgk_time_stages", "dgk_num_variants", "dgk_set_variant", "dgk_advect3d_eq_cal_tend", "dgk_divlike_dirXYZ_batched",
    "dgk_lift_hexahedral_batched", "dgk_ensemble_create", "dgk_ensemble_destroy",
    "dgk_ensemble_set_fields", "dgk_ensemble_get_field", "dgk_ensemble_set_dt",
    "dgk_ensemble_step", "dgk_ensemble_run", "dgk_ensemble_time_stages",
    "dgk_ensemble_tune", "dgk_ensemble_query", "dgk_ensemble_set_shape", "dgk_autotune", "dgk_diagnostics",
    "dgk_diagnostics_noalloc",
    "dgk_ensemble_diagnostics",
    "dgk_matvec_dir_batched", "dgk_gradlike_dirXYZ_batched", "dgk_divlike_dirXYZ_var5_batched",
    "dgk_mesh_setup_device",
)


class DGKError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dgk error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Loads libdgk.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with "
                          "`make -C scale-dg-kernel-extraction_amd/csrc` or "
                          "`python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(LIB_PATH)
    i, d, vp, ll = C.c_int, C.c_double, C.c_void_p, C.c_longlong
    sig = {
        "dgk_last_error": ([], C.c_char_p),
        "dgk_version": ([], i),
        "dgk_optr_kernel_setup": ([C.c_char_p, vp], i),
        "dgk_read_operator": ([C.c_char_p, i, vp, vp, vp], i),
        "dgk_mesh_sizes": ([i, i, i, i, vp], i),
        "dgk_mesh_setup": ([i, i, i, i, d, d, d] + [vp] * 8, i),
        "dgk_create": ([vp, i, i, i, i] + [vp] * 8 + [i, i, i], i),
        "dgk_create_structured": ([vp, i, i, i, i, d, d, d, vp, vp, i, i], i),
        "dgk_destroy": ([vp], i),
        "dgk_set_fields": ([vp] * 5, i),
        "dgk_get_field": ([vp, i, vp], i),
        "dgk_field_ptr": ([vp, i, vp], i),
        "dgk_cal_tend": ([vp], i),
        "dgk_rk_stage": ([vp, i, d], i),
        "dgk_step": ([vp, d], i),
        "dgk_run": ([vp, d, i, vp], i),
        "dgk_synchronize": ([vp], i),
        "dgk_query": ([vp, i, vp], i),
        "dgk_time_stages": ([vp, d, i, vp], i),
        "dgk_num_variants": ([vp, vp], i),
        "dgk_set_variant": ([vp, i], i),
        "dgk_advect3d_eq_cal_tend": ([vp] * 13 + [i] * 6, i),
        "dgk_divlike_dirXYZ_batched": ([i, i, i] + [vp] * 9, i),
        "dgk_lift_hexahedral_batched": ([i, i, vp, vp, vp, vp], i),
        "dgk_autotune": ([vp, d, i, vp, vp], i),
        "dgk_mesh_setup_device": ([i, i, i, i, d, d, d] + [vp] * 9, i),
        "dgk_matvec_dir_batched": ([i, i, i, vp, vp, vp, vp], i),
        "dgk_gradlike_dirXYZ_batched": ([i, i] + [vp] * 7, i),
        "dgk_divlike_dirXYZ_var5_batched": ([i, i] + [vp] * 9, i),
        "dgk_diagnostics": ([vp, vp, vp, vp], i),
        "dgk_diagnostics_noalloc": ([vp, vp, vp, vp], i),
        "dgk_ensemble_create": ([vp, i] + [vp] * 5 + [C.c_char_p, i, i], i),
        "dgk_ensemble_destroy": ([vp], i),
        "dgk_ensemble_set_fields": ([vp, i] + [vp] * 4, i),
        "dgk_ensemble_get_field": ([vp, i, vp], i),
        "dgk_ensemble_set_dt": ([vp, vp], i),
        "dgk_ensemble_step": ([vp], i),
        "dgk_ensemble_run": ([vp, i], i),
        "dgk_ensemble_time_stages": ([vp, i, vp], i),
        "dgk_ensemble_tune": ([vp, i, vp], i),
        "dgk_ensemble_query": ([vp, i, vp], i),
        "dgk_ensemble_set_shape": ([vp, i, vp], i),
        "dgk_ensemble_diagnostics": ([vp, i, vp, vp, vp], i),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(rc):
    if rc != OK:
        raise DGKError(rc, lib().dgk_last_error().decode(errors="replace"))


def _p(a, dtype=np.float64):
    """Address of a host array that the library reads or writes as a
    contiguous column-major block of dtype (double, or int32_t for indices
    and sizes).  Other element types and strided views are refused rather
    than converted: the library takes the bytes as they lie."""
    if a is None:
        return None
    if a.dtype != dtype:
        raise TypeError(f"array of {a.dtype} where the library expects {np.dtype(dtype)}")
    if not (a.flags.f_contiguous if a.ndim > 1 else a.flags.c_contiguous):
        raise ValueError("array must be contiguous in Fortran order")
    return a.ctypes.data_as(C.c_void_p)


def optr_kernel_setup(evaltype: str) -> int:
    """DGOptrKernel_OptType name -> type id (GENERAL 1, OPT1 2, HIP 3)."""
    t = C.c_int(0)
    _check(lib().dgk_optr_kernel_setup(evaltype.encode(), C.byref(t)))
    return t.value


def read_operator(p: int, opdir: str = OPERATOR_DIR):
    """x1D(Nq), D1D(Nq,Nq), Lift_mat(Nq,Nq,Nq,6) from operator_data/p<p>.dat."""
    Nq = p + 1
    x1D = np.zeros(Nq)
    D = np.zeros((Nq, Nq), order="F")
    L = np.zeros((Nq, Nq, Nq, 6), order="F")
    _check(lib().dgk_read_operator(opdir.encode(), p, _p(x1D), _p(D), _p(L)))
    return x1D, D, L


def mesh_sizes(NeX, NeY, NeZ, p):
    s = np.zeros(7, dtype=np.int32)
    _check(lib().dgk_mesh_sizes(NeX, NeY, NeZ, p, _p(s, np.int32)))
    return dict(zip(("Nq", "Np", "Nfp", "NfpTot", "Ne", "NeA", "NhaloNode"),
                    (int(v) for v in s)))


def mesh_setup(NeX, NeY, NeZ, p, dom=(1.0, 1.0, 1.0), x1D=None, tables=True):
    """Reference-layout mesh tables (mod_mesh.f90:74-193) as a dict."""
    sz = mesh_sizes(NeX, NeY, NeZ, p)
    Np, Nf, Ne = sz["Np"], sz["NfpTot"], sz["Ne"]
    out = dict(sz)
    if x1D is None:
        x1D = read_operator(p)[0]
    pos = np.zeros((Np, Ne, 3), order="F")
    halo = np.zeros(sz["NhaloNode"], dtype=np.int32)
    arrs = dict(VMapM=None, VMapP=None, normal_fn=None, Escale=None, Fscale=None)
    if tables:
        arrs = dict(VMapM=np.zeros((Nf, Ne), dtype=np.int32, order="F"),
                    VMapP=np.zeros((Nf, Ne), dtype=np.int32, order="F"),
                    normal_fn=np.zeros((Nf, Ne, 3), order="F"),
                    Escale=np.zeros((Np, Ne, 3), order="F"),
                    Fscale=np.zeros((Nf, Ne), order="F"))
    _check(lib().dgk_mesh_setup(NeX, NeY, NeZ, p, float(dom[0]), float(dom[1]),
                                float(dom[2]), _p(x1D), _p(arrs["VMapM"], np.int32),
                                _p(arrs["VMapP"], np.int32), _p(arrs["normal_fn"]),
                                _p(arrs["Escale"]), _p(arrs["Fscale"]), _p(pos),
                                _p(halo, np.int32)))
    out.update(arrs)
    out.update(pos_en=pos, halo_src_map=halo, x1D=x1D)
    return out


class Solver:
    """Device-resident tendency/RK solver (dgk_create / dgk_create_structured)."""

    def __init__(self, handle, Np, Ne, NeA):
        self._h = handle
        self.Np, self.Ne, self.NeA = Np, Ne, NeA

    @classmethod
    def from_tables(cls, Nq, Ne, NeA, D1D, Lift, VMapM, VMapP, normal_fn, Escale,
                    Fscale, halo_src_map, typeid=TYPE_HIP, flags=0):
        h = C.c_void_p()
        D1D, Lift = np.asfortranarray(D1D), np.asfortranarray(Lift)
        _check(lib().dgk_create(C.byref(h), Nq, Ne, NeA, 6 * Nq * Nq, _p(D1D), _p(Lift),
                                _p(VMapM, np.int32), _p(VMapP, np.int32), _p(normal_fn),
                                _p(Escale), _p(Fscale), _p(halo_src_map, np.int32),
                                len(halo_src_map),
                                typeid, flags))
        return cls(h, Nq ** 3, Ne, NeA)

    @classmethod
    def structured(cls, NeX, NeY, NeZ, p, D1D, Lift, dom=(1.0, 1.0, 1.0),
                   typeid=TYPE_HIP, flags=0):
        h = C.c_void_p()
        D1D, Lift = np.asfortranarray(D1D), np.asfortranarray(Lift)
        _check(lib().dgk_create_structured(C.byref(h), NeX, NeY, NeZ, p, float(dom[0]),
                                           float(dom[1]), float(dom[2]), _p(D1D),
                                           _p(Lift), typeid, flags))
        sz = mesh_sizes(NeX, NeY, NeZ, p)
        return cls(h, sz["Np"], sz["Ne"], sz["NeA"])

    def close(self):
        if self._h:
            lib().dgk_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _full(self, a):
        """Accepts (Np,Ne) or (Np,NeA) and returns a (Np,NeA) F-array."""
        a = np.asarray(a, dtype=np.float64)
        if a.shape == (self.Np, self.NeA):
            return np.asfortranarray(a)
        if a.shape != (self.Np, self.Ne):
            raise ValueError(f"field shape {a.shape} is neither (Np,Ne) nor (Np,NeA)")
        f = np.zeros((self.Np, self.NeA), order="F")
        f[:, :self.Ne] = a
        return f

    def set_fields(self, q, u, v, w):
        q, u, v, w = (self._full(a) for a in (q, u, v, w))
        _check(lib().dgk_set_fields(self._h, _p(q), _p(u), _p(v), _p(w)))

    def get_field(self, which=FIELD_Q, owned=True):
        out = np.zeros((self.Np, self.NeA), order="F")
        _check(lib().dgk_get_field(self._h, which, _p(out)))
        return out[:, :self.Ne] if owned else out

    def field_ptr(self, which=FIELD_Q):
        p = C.c_void_p()
        _check(lib().dgk_field_ptr(self._h, which, C.byref(p)))
        return p.value

    def cal_tend(self):
        _check(lib().dgk_cal_tend(self._h))
        return self.get_field(FIELD_DQDT)

    def rk_stage(self, stage, dt):
        _check(lib().dgk_rk_stage(self._h, stage, float(dt)))

    def step(self, dt):
        _check(lib().dgk_step(self._h, float(dt)))

    def run(self, dt, nstep, minmax=False):
        mm = np.zeros((2, max(nstep, 1)), order="F") if minmax else None
        _check(lib().dgk_run(self._h, float(dt), nstep, _p(mm)))
        return mm[:, :nstep] if minmax else None

    def synchronize(self):
        _check(lib().dgk_synchronize(self._h))

    def query(self, what):
        v = C.c_int64(0)
        _check(lib().dgk_query(self._h, what, C.byref(v)))
        return v.value

    @property
    def compressed(self):
        return bool(self.query(0))

    @property
    def kernel_kind(self):
        return "pencil" if self.query(6) else "node"

    @property
    def num_variants(self):
        n = C.c_int(0)
        _check(lib().dgk_num_variants(self._h, C.byref(n)))
        return n.value

    @property
    def variant(self):
        return self.query(8)

    def set_variant(self, v):
        _check(lib().dgk_set_variant(self._h, int(v)))

    def time_stages(self, dt, nstage):
        ms = C.c_double(0.0)
        _check(lib().dgk_time_stages(self._h, float(dt), nstage, C.byref(ms)))
        return ms.value

    def autotune(self, dt, nrep=10):
        """Selects the fastest tuning variant; returns (variant, ms per stage)."""
        v, ms = C.c_int(0), C.c_double(0.0)
        _check(lib().dgk_autotune(self._h, float(dt), nrep, C.byref(v), C.byref(ms)))
        return v.value, ms.value

    def diagnostics(self, x1D, ref=None):
        """Device reductions of the state: min, max, LGL mass, |q|-mass, sum
        q^2 and, against ref (another Solver or a device address of a (Np,Ne)
        field), rel_l2 and max_err."""
        w = lgl_weights(x1D)
        out = np.zeros(8)
        rp = ref.field_ptr() if isinstance(ref, Solver) else ref
        _check(lib().dgk_diagnostics(self._h, _p(w), rp, _p(out)))
        d = dict(min=out[0], max=out[1], mass=out[2], abs_mass=out[3], sum_sq=out[4])
        if rp:
            d.update(rel_l2=float(np.sqrt(out[5] / out[6])) if out[6] > 0 else float("nan"),
                     max_err=out[7], err_sq=out[5], ref_sq=out[6])
        return d


def lgl_weights(x1D):
    """Quadrature weights of the Legendre-Gauss-Lobatto nodes x1D:
    w_i = 2 / (p (p+1) P_p(x_i)^2)."""
    x1D = np.asarray(x1D, dtype=np.float64)
    p = len(x1D) - 1
    P = np.polynomial.legendre.Legendre.basis(p)(x1D)
    return np.ascontiguousarray(2.0 / (p * (p + 1) * P * P))


def advect3d_eq_cal_tend(q, u, v, w, D1D, D1D_tr, Lift, VMapM, VMapP, normal_fn,
                         Escale, Fscale, Nq, Np, NfpTot, Ne, NeA, typeid=TYPE_OPT1):
    """Stateless drop-in of advect3d_eq_cal_tend on host arrays (q,u,v,w with
    halos, shape (Np,NeA)); returns dqdt(Np,Ne)."""
    dqdt = np.zeros((Np, Ne), order="F")
    arrs = [np.asfortranarray(a) for a in (q, u, v, w, D1D, D1D_tr, Lift)]
    _check(lib().dgk_advect3d_eq_cal_tend(_p(dqdt), *[_p(a) for a in arrs],
                                          _p(VMapM, np.int32), _p(VMapP, np.int32),
                                          _p(normal_fn), _p(Escale),
                                          _p(Fscale), Nq, Np, NfpTot, Ne, NeA, typeid))
    return dqdt


def divlike_batched(Nq, nelem, typeid, Mat, Mat_tr, in_x, in_y, in_z, out_x, out_y,
                    out_z, stream=None):
    """Device pointers (ints) for every array argument."""
    _check(lib().dgk_divlike_dirXYZ_batched(Nq, nelem, typeid, Mat, Mat_tr, in_x, in_y,
                                            in_z, out_x, out_y, out_z, stream))


def lift_batched(Nq, nelem, Lift, fin, fout, stream=None):
    _check(lib().dgk_lift_hexahedral_batched(Nq, nelem, Lift, fin, fout, stream))


def mesh_setup_device(NeX, NeY, NeZ, p, x1D, ptrs, dom=(1.0, 1.0, 1.0), stream=None):
    """dgk_mesh_setup_device: ptrs maps VMapM, VMapP, normal_fn, Escale, Fscale,
    pos_en, halo_src_map to device addresses (missing keys: not built)."""
    x1D = np.ascontiguousarray(x1D, dtype=np.float64)
    keys = ("VMapM", "VMapP", "normal_fn", "Escale", "Fscale", "pos_en", "halo_src_map")
    _check(lib().dgk_mesh_setup_device(NeX, NeY, NeZ, p, float(dom[0]), float(dom[1]),
                                       float(dom[2]), _p(x1D), *(ptrs.get(k) for k in keys),
                                       stream))


def matvec_dir_batched(Nq, nelem, direction, Mat, fin, fout, stream=None):
    """dirX/dirY/dirZ (direction 0/1/2); device pointers."""
    _check(lib().dgk_matvec_dir_batched(Nq, nelem, direction, Mat, fin, fout, stream))


def gradlike_batched(Nq, nelem, Mat, Mat_tr, fin, out_x, out_y, out_z, stream=None):
    _check(lib().dgk_gradlike_dirXYZ_batched(Nq, nelem, Mat, Mat_tr, fin, out_x, out_y,
                                             out_z, stream))


def divlike_var5_batched(Nq, nelem, Mat, Mat_tr, in_x, in_y, in_z, out_x, out_y, out_z,
                         stream=None):
    _check(lib().dgk_divlike_dirXYZ_var5_batched(Nq, nelem, Mat, Mat_tr, in_x, in_y, in_z,
                                                 out_x, out_y, out_z, stream))


class Ensemble:
    """Many independent periodic runs advanced together (dgk_ensemble_*).

    ``specs`` is a sequence of dicts with keys ``p`` and ``ne`` = (NeX,NeY,NeZ)
    and optionally ``dom`` (domain lengths, default unit cube).  Every
    instance evolves exactly as a ``Solver.structured`` run with the same
    flags in OPT1's summation order."""

    def __init__(self, specs, flags=0, typeid=TYPE_HIP, opdir: str = OPERATOR_DIR):
        self.specs = [dict(s) for s in specs]
        n = len(self.specs)
        p = np.array([s["p"] for s in self.specs], dtype=np.int32)
        ne = np.array([s["ne"] for s in self.specs], dtype=np.int32).reshape(n, 3)
        nex, ney, nez = (np.ascontiguousarray(ne[:, d]) for d in range(3))
        dom = np.array([s.get("dom", (1.0, 1.0, 1.0)) for s in self.specs],
                       dtype=np.float64).reshape(n, 3)
        dom = np.ascontiguousarray(dom).ravel()  # n triples
        h = C.c_void_p()
        _check(lib().dgk_ensemble_create(C.byref(h), n, _p(p, np.int32), _p(nex, np.int32),
                                         _p(ney, np.int32), _p(nez, np.int32),
                                         _p(dom), opdir.encode(), typeid, flags))
        self._h = h
        self.n = n
        self.shapes = [((s["p"] + 1) ** 3, int(np.prod(s["ne"]))) for s in self.specs]

    def close(self):
        if self._h:
            lib().dgk_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fields(self, i, q, u, v, w):
        arrs = [np.asfortranarray(a, dtype=np.float64) for a in (q, u, v, w)]
        for a in arrs:
            if a.shape != self.shapes[i]:
                raise ValueError(f"instance {i}: expected shape {self.shapes[i]}, got {a.shape}")
        _check(lib().dgk_ensemble_set_fields(self._h, i, *[_p(a) for a in arrs]))

    def get_field(self, i):
        q = np.empty(self.shapes[i], order="F")
        _check(lib().dgk_ensemble_get_field(self._h, i, _p(q)))
        return q

    def set_dt(self, dt):
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        if dt.shape != (self.n,):
            raise ValueError("one dt per instance")
        _check(lib().dgk_ensemble_set_dt(self._h, _p(dt)))

    def step(self):
        _check(lib().dgk_ensemble_step(self._h))

    def run(self, nstep):
        _check(lib().dgk_ensemble_run(self._h, nstep))

    def time_stages(self, nstage):
        ms = C.c_double(0.0)
        _check(lib().dgk_ensemble_time_stages(self._h, nstage, C.byref(ms)))
        return ms.value

    def tune(self, nrep=3):
        """Per-class stage time in ms (classes by increasing order)."""
        ms = np.zeros(self.query(4))
        _check(lib().dgk_ensemble_tune(self._h, nrep, _p(ms)))
        return ms

    def query(self, what):
        v = C.c_longlong(0)
        _check(lib().dgk_ensemble_query(self._h, what, C.byref(v)))
        return v.value

    def set_shape(self, threads=0, elems_per_item=None):
        """Tile shape by hand: threads per workgroup (192, 256; 0 default)
        and elements per item of each class (0 or None: as many as fit)."""
        e = None
        if elems_per_item is not None:
            e = np.ascontiguousarray(elems_per_item, dtype=np.int32)
            if e.shape != (self.query(4),):
                raise ValueError("one entry per order class")
        _check(lib().dgk_ensemble_set_shape(self._h, threads, None if e is None else _p(e, np.int32)))

    @property
    def shape(self):
        """(threads per workgroup, [elements per item of each class])."""
        return self.query(6), [self.query(48 + c) for c in range(self.query(4))]

    def diagnostics(self, i, x1D, ref=None):
        """Solver.diagnostics for instance i (ref: a device address of a
        (Np,Ne) field, or None)."""
        w = lgl_weights(x1D)
        out = np.zeros(8)
        _check(lib().dgk_ensemble_diagnostics(self._h, i, _p(w), ref, _p(out)))
        d = dict(min=out[0], max=out[1], mass=out[2], abs_mass=out[3], sum_sq=out[4])
        if ref:
            d.update(rel_l2=float(np.sqrt(out[5] / out[6])) if out[6] > 0 else float("nan"),
                     max_err=out[7], err_sq=out[5], ref_sq=out[6])
        return d

    @property
    def classes(self):
        """[(Nq, workgroups)] of the persistent grid."""
        return [(self.query(16 + c), self.query(32 + c)) for c in range(self.query(4))]
