"""dgk_diagnostics_noalloc: the diagnostics reductions with their partial
sums in the library's file-scope device array instead of an allocation.
Checked through the C-ABI against dgk_diagnostics on the same state, bit for
bit (same kernels, same reduction order), on two calls in a row, the second
on a smaller field, with and without a reference field."""
import ctypes as C

import numpy as np
import pytest

import dgk
from conftest import say
from dgk import ic

pytestmark = pytest.mark.gpu


def _solver(ne, p, seed, flags=0):
    x1D, D, L = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p, x1D=x1D, tables=False)
    q = ic.sine_modes(m["pos_en"], seed)
    s = dgk.Solver.structured(*ne, p, D, L, flags=flags)
    s.set_fields(q, *(np.full_like(q, c) for c in (0.3, -0.2, 0.5)))
    return s, x1D


def _noalloc(torch, s, x1D, ref):
    w = dgk.lgl_weights(x1D)
    out = torch.full((8,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rp = ref.field_ptr() if ref is not None else None
    rc = dgk.lib().dgk_diagnostics_noalloc(s._h, w.ctypes.data_as(C.c_void_p), rp,
                                           C.c_void_p(out.data_ptr()))
    assert rc == 0, dgk.lib().dgk_last_error()
    s.synchronize()
    return out.cpu().numpy()


def _alloc(s, x1D, ref):
    w = dgk.lgl_weights(x1D)
    out = np.zeros(8)
    rp = ref.field_ptr() if ref is not None else None
    assert dgk.lib().dgk_diagnostics(s._h, w.ctypes.data_as(C.c_void_p), rp,
                                     out.ctypes.data_as(C.c_void_p)) == 0
    return out


def test_noalloc_diagnostics_equal_allocating(gpu):
    torch = gpu
    # the first field needs all 1024 partial blocks, the second 8
    cases = [((24, 20, 16), 7, 71), ((4, 4, 2), 3, 72)]
    for ne, p, seed in cases:
        say(f"[noalloc] {ne} p={p}")
        s, x1D = _solver(ne, p, seed)
        s.run(ic.cfl_dt(p, (0.3, -0.2, 0.5), ne), 2)
        r, _ = _solver(ne, p, seed + 1)
        for ref in (None, r):
            got = _noalloc(torch, s, x1D, ref)
            want = _alloc(s, x1D, ref)
            assert np.array_equal(got, want), (ne, got, want)
        s.close()
        r.close()


def test_noalloc_diagnostics_errors(gpu):
    s, x1D = _solver((2, 2, 2), 1, 5)
    w = dgk.lgl_weights(x1D)
    rc = dgk.lib().dgk_diagnostics_noalloc(s._h, w.ctypes.data_as(C.c_void_p), None, None)
    assert rc == dgk.ERR_ARG
    s.close()
