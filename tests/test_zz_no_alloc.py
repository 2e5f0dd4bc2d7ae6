"""sphgpu_timestep_noalloc (the entry point that allocates no device memory)
against sphgpu_timestep on the same particles: two calls in a row, the
second on a smaller state, each bit-equal (dt is a minimum of the same
per-particle quotients, so the order of the particles does not matter).
Also: its refusals, and that it leaves the caller's arrays untouched."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from sphcode_amd import ics

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


def _bind():
    from sphcode_amd import library
    lib, _ = library()
    f = lib.sphgpu_timestep_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_double, C.c_double, C.c_void_p, C.c_void_p]
    return lib, f


def _reference_dt(p, f):
    """sphgpu_timestep on an initialised context, with the state it read."""
    from sphcode_amd import GPUEngine
    g = GPUEngine(p)
    g.upload(f)
    g.initialize()
    dt = g.timestep()
    st = g.download(("sml", "acc"))
    hvs = g.time()[2]
    g.close()
    return dt, st, hvs


def _noalloc_dt(fn, p, st, hvs):
    import torch
    dev = torch.device("cuda:0")
    sml = torch.tensor(st["sml"], dtype=torch.float64, device=dev)
    acc = [torch.tensor(np.ascontiguousarray(st["acc"][:, k]), dtype=torch.float64, device=dev) for k in range(p.dim)]
    hv = torch.tensor([hvs], dtype=torch.float64, device=dev)
    out = torch.full((1,), -1.0, dtype=torch.float64, device=dev)
    before = [a.clone() for a in acc] + [sml.clone()]
    stream = torch.cuda.current_stream(dev).cuda_stream
    ptr = [a.data_ptr() for a in acc] + [0] * (3 - p.dim)
    rc = fn(int(p.dim), len(st["sml"]), sml.data_ptr(), ptr[0], ptr[1], ptr[2], hv.data_ptr(),
            float(p.cfl_sound), float(p.cfl_force), out.data_ptr(), C.c_void_p(stream))
    assert rc == 0
    torch.cuda.synchronize(dev)
    for x, y in zip(before, acc + [sml]):
        assert torch.equal(x, y)
    return float(out.item())


def test_noalloc_timestep_matches_timestep(gpu):
    lib, fn = _bind()
    cases = [ics.evrard(4096, leaf=32), ics.evrard(1024, leaf=8)]
    for k, (p, f) in enumerate(cases):
        dt_ref, st, hvs = _reference_dt(p, f)
        dt = _noalloc_dt(fn, p, st, hvs)
        print(f"[noalloc] call {k + 1}: n={len(f['mass'])} dt {dt!r} sphgpu_timestep {dt_ref!r}", flush=True)
        assert dt == dt_ref
    # a 2D state, then a 1D one, on the same scratch array
    for p, f in (ics.khi(n_low=32, n_high=45, jitter=0.05), ics.sod(400, 50)):
        dt_ref, st, hvs = _reference_dt(p, f)
        assert _noalloc_dt(fn, p, st, hvs) == dt_ref


def test_noalloc_timestep_refusals(gpu):
    lib, fn = _bind()
    import torch
    x = torch.zeros(8, dtype=torch.float64, device="cuda:0")
    p = x.data_ptr()
    assert fn(4, 8, p, p, p, p, p, 0.3, 0.125, p, None) == -1
    assert fn(3, 0, p, p, p, p, p, 0.3, 0.125, p, None) == -1
    assert fn(3, (1 << 30) + 1, p, p, p, p, p, 0.3, 0.125, p, None) == -1
    assert fn(3, 8, p, p, p, None, p, 0.3, 0.125, p, None) == -1
    lib.sphgpu_last_error.restype = C.c_char_p
    assert b"n must be in" in lib.sphgpu_last_error() or b"null" in lib.sphgpu_last_error()
