"""sphx_update_cell_linked_list_noalloc (the allocation-free cell-list update,
workspace in a fixed device array of the library) against
sphx_update_cell_linked_list: bit-identical cell offsets and particle indices
on two calls in a row, the second on a smaller body (particles in random
order, then in cell order), and the capacity check at the cell limit."""
import ctypes as C

import numpy as np
import pytest

from sphinxsys_amd import _capi as capi
from sphinxsys_amd import ck

pytestmark = pytest.mark.gpu

MAX_CELLS = 1 << 22  # the workspace limit stated in include/sphx.h


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
    # first call: 300k particles in random order, 3D
    big = rng.uniform(-0.1, 1.1, size=(300000, 3))
    print("no-alloc cell list: 300000 random 3D particles", flush=True)
    check_case(hip, oracle, big, 0.004)
    # second call, smaller: a 2D lattice in cell order
    small = lattice(2, 70, 0.0125)
    print("no-alloc cell list: 4900 lattice 2D particles", flush=True)
    check_case(hip, oracle, small, 0.0125)


def test_noalloc_at_the_cell_limit_and_beyond(hip):
    """A mesh of exactly the workspace's cell count works and equals the
    allocating update; one more row of cells is refused before any launch."""
    rng = np.random.default_rng(7)
    dim, dp = 2, 0.0025
    gs = 2.0 * 1.3 * dp
    # the mesh spans ceil(extent / gs) + 4 cells per axis (2-cell buffer each side)
    per_axis = int(MAX_CELLS ** 0.5)  # 2048 x 2048 = the limit exactly
    for cells_y, ok in ((per_axis, True), (per_axis + 1, False)):
        ext_x, ext_y = (per_axis - 4) * gs, (cells_y - 4) * gs
        pos = rng.uniform((0.0, 0.0), (ext_x, ext_y), size=(200000, dim))
        ctx = ck.Context(hip)
        b = ck.Body(ctx, dim, len(pos), dp)
        b.write("Position", pos)
        c = ck.CellLinkedList(b, np.zeros(dim), np.array([ext_x, ext_y]), grid_spacing=gs)
        try:
            assert (c.n_cells <= MAX_CELLS) == ok, (c.all_cells, c.n_cells)
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
