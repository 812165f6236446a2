"""sphx_update_cell_linked_list_noalloc (the allocation-free cell-list update,
workspace in a fixed device array of the library) against
sphx_update_cell_linked_list: bit-identical cell offsets and particle indices
on two calls in a row, the second on a smaller body, both paths of the run
placement (many runs: the radix passes; few runs: the one-workgroup sort),
and the capacity check."""
import ctypes as C

import numpy as np
import pytest

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import ck

pytestmark = pytest.mark.gpu

MAX_PARTICLES = 1 << 20  # the workspace limit stated in include/sphx.h


def lattice(dim, m, dp):
    axes = [(np.arange(m) + 0.5) * dp for _ in range(dim)]
    return np.stack([g.ravel() for g in np.meshgrid(*axes, indexing="ij")], axis=1)


def both_updates(lib, pos, dp):
    """Cell list of `pos` by the allocation-free entry point, then by the
    allocating one, on one body and cell list."""
    dim = pos.shape[1]
    ctx = ck.Context(lib)
    b = ck.Body(ctx, dim, len(pos), dp)
    b.write("Position", pos)
    c = ck.CellLinkedList(b, np.zeros(dim), np.ones(dim))
    try:
        lib.call("sphx_update_cell_linked_list_noalloc", c.h)
        no_alloc = c.read()
        lib.call("sphx_update_cell_linked_list", c.h)
        alloc = c.read()
        # and the allocation-free one again over the allocating one's result
        lib.call("sphx_update_cell_linked_list_noalloc", c.h)
        again = c.read()
    finally:
        c.close(), b.close(), ctx.close()
    return no_alloc, alloc, again


def check_case(hip, oracle, pos, dp):
    (o1, i1), (o2, i2), (o3, i3) = both_updates(hip, pos, dp)
    assert np.array_equal(o1, o2) and np.array_equal(i1, i2)
    assert np.array_equal(o3, o2) and np.array_equal(i3, i2)
    # and the oracle's canonical cell list of the same positions
    (oo, oi), _, _ = both_updates(oracle, pos, dp)
    assert np.array_equal(o1, oo) and np.array_equal(i1, oi)
    assert o1[-1] == len(pos)


def test_noalloc_matches_allocating_update_large_then_small(hip, oracle):
    rng = np.random.default_rng(20261203)
    # first call: 300k particles in random order (as many runs as particles:
    # the radix passes), 3D
    big = rng.uniform(-0.1, 1.1, size=(300000, 3))
    print("no-alloc cell list: 300000 random 3D particles", flush=True)
    check_case(hip, oracle, big, 0.004)
    # second call, smaller: a 2D lattice in cell order (few runs: one workgroup)
    small = lattice(2, 70, 0.0125)
    print("no-alloc cell list: 4900 lattice 2D particles", flush=True)
    check_case(hip, oracle, small, 0.0125)


def test_noalloc_at_the_particle_limit_and_beyond(hip):
    rng = np.random.default_rng(7)
    dim = 2
    for n, ok in ((MAX_PARTICLES, True), (MAX_PARTICLES + 1, False)):
        pos = rng.uniform(0.0, 1.0, size=(n, dim))
        ctx = ck.Context(hip)
        b = ck.Body(ctx, dim, n, 0.002)
        b.write("Position", pos)
        c = ck.CellLinkedList(b, np.zeros(dim), np.ones(dim))
        try:
            status = hip.lib.sphx_update_cell_linked_list_noalloc(c.h)
            if ok:
                assert status == capi.SPHX_OK, hip.lib.sphx_last_error().decode()
                off, idx = c.read()
                hip.call("sphx_update_cell_linked_list", c.h)
                off2, idx2 = c.read()
                assert np.array_equal(off, off2) and np.array_equal(idx, idx2)
            else:
                assert status == 3  # SPHX_E_CAPACITY, before any launch
                assert b"exceed its workspace" in hip.lib.sphx_last_error()
        finally:
            c.close(), b.close(), ctx.close()
