"""trisk_diagnostics_noalloc (the allocation-free, stream-ordered diagnostics for callers that
cannot allocate) against trisk_diagnostics, through the C ABI: bitwise equal sums on two
calls in a row, the second on a smaller mesh, for the TRiSK model in both precisions and
for the ICON triangle model."""
import ctypes as C
import sys

import numpy as np
import pytest

from dmplex_amd import HIPBackend, ICONShallowWater, TRiSKShallowWater, create_hexagon_mesh, create_spherical_mesh
from dmplex_amd import _device as D
from dmplex_amd.synthetic import earth_parameters, seeded_state

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _stream():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _both(m, y_host):
    """(no-alloc sums, trisk_diagnostics sums) of the host state y_host."""
    import torch

    y = m.upload_state(y_host)
    out = torch.full((3,), float("nan"), dtype=torch.float64, device=y.device)
    D.check(D.lib().trisk_diagnostics_noalloc(m.handle, C.c_void_p(y.data_ptr()), C.c_void_p(out.data_ptr()),
                                              _stream()))
    torch.cuda.current_stream().synchronize()
    got = out.cpu().numpy()
    ref = np.zeros(3)
    D.check(D.lib().trisk_diagnostics(m.handle, C.c_void_p(y.data_ptr()), ref.ctypes.data_as(C.POINTER(C.c_double)),
                                      _stream()))
    return got, ref


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
def test_noalloc_diagnostics_equal_trisk_diagnostics(precision):
    for level in (5, 3):  # the second call on a smaller mesh
        say(f"no-alloc diagnostics: TRiSK level {level} {precision}")
        dm = create_hexagon_mesh(level, centroidal=True)
        m = TRiSKShallowWater(dm, centers=dm.generators, backend=HIPBackend(0, precision), **earth_parameters())
        got, ref = _both(m, seeded_state(m.grid))
        assert np.all(np.isfinite(ref))
        assert np.array_equal(got, ref), (level, got, ref)


def test_noalloc_diagnostics_icon_and_bad_pointer():
    for level in (4, 2):
        say(f"no-alloc diagnostics: ICON level {level}")
        m = ICONShallowWater(create_spherical_mesh(level), **earth_parameters())
        got, ref = _both(m, seeded_state(m.grid))
        assert np.array_equal(got, ref), (level, got, ref)
    rc = D.lib().trisk_diagnostics_noalloc(m.handle, None, None, _stream())
    assert rc == -1 and "null pointer" in D.lib().trisk_last_error().decode()
