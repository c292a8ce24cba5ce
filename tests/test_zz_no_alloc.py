"""jm_block_bcgstab_noalloc, the block BiCGStab on the library's static
per-device work space, against jm_block_bcgstab (the same kernels on a work
space allocated for the call): bit-identical corrections, on two calls in a
row whose second system is smaller than the first, so whatever the first
call left in the shared array must not reach the second. Both are also held
to the reference's own correction at the bar of the context's solve (1e-9
relative, tests/test_gpu_block_solver.py).

The size limit is checked on the host before any device call (a CPU test in
tests/test_abi.py)."""
import numpy as np
import pytest

from myjoe_amd import joemi
from tests.helpers import compare
from tests.test_gpu_block_solver import reference_system

pytestmark = pytest.mark.gpu


def _solve(fn, fixture):
    ni, nv, A, rhs, dq = reference_system(fixture)
    t = joemi.DeviceArray
    phi = t(np.zeros(rhs.shape))
    it, st = fn(t(ni), t(nv), t(A), t(rhs), phi)
    return it, st, phi.numpy(), dq, len(ni) - 1


def test_noalloc_equals_allocating_solve_on_two_calls():
    # box4 (224 cells), then the smaller bump24 (192 cells)
    sizes = []
    for fixture in ("box4_hllc", "bump24_hllc"):
        print(f"noalloc {fixture}", flush=True)
        it_a, st_a, got_a, dq, ncv = _solve(joemi.block_bcgstab, fixture)
        it_n, st_n, got_n, _, _ = _solve(joemi.block_bcgstab_noalloc, fixture)
        sizes.append(ncv)
        assert (it_n, st_n) == (it_a, st_a), fixture
        assert np.array_equal(got_n, got_a), fixture
        compare(got_n, dq, 1e-9, f"{fixture}:DQ")
    assert sizes[1] < sizes[0]
    assert joemi.block_bcgstab_noalloc_max_cells() >= 100_000
