"""Host-side case geometry: boxes, add/subtract complex shapes, lattice particle
generation and wall normals.

Restates the subset of SPHinXsys geometry the dam-break cases use:
  - GeometricShapeBox / TransformShape<GeometricBox> contain and closest point
    (src/shared/geometries/geometric_element.cpp:19-59, geometric_element.h:43-55,
    transform_geometry.h:53-64, common/geometric_primitive.h:193-208);
  - ComplexShape add/subtract (geometries/base_geometry.cpp:84-137);
  - ParticleGenerator<BaseParticles, Lattice>
    (src/for_3D_build/particle_generator/particle_generator_lattice_3d.cpp:12-25,
     src/for_2D_build/particle_generator/particle_generator_lattice_2d.cpp:12-23)
    over the system bounds expanded by 4 dp (sphinxsys_system/sph_system.cpp:39);
  - Shape::findNormalDirection (geometries/base_geometry.cpp:45-58) as used by
    NormalFromBodyShapeCK (shared_ck/.../general_dynamics/geometric_dynamics.cpp:19-27).
All arithmetic is IEEE double, elementwise, in the reference's operation order.
"""
import math

import numpy as np


class Box:
    """Axis-aligned box with centre `translation` and `halfsize`."""

    def __init__(self, translation, halfsize):
        self.t = np.asarray(translation, dtype=np.float64)
        self.h = np.asarray(halfsize, dtype=np.float64)

    def contains(self, p):
        # GeometricBox::checkContain: outside if |x_k| > halfsize_k on any axis
        local = p - self.t
        return np.all(np.abs(local) <= self.h, axis=-1)

    def closest_point(self, p):
        # GeometricBox::findClosestPoint (Simbody Geo_Box algorithm)
        c = p - self.t
        inside = np.ones(c.shape[0], dtype=bool)
        for k in range(c.shape[1]):
            lo = c[:, k] < -self.h[k]
            hi = c[:, k] > self.h[k]
            c[lo, k] = -self.h[k]
            c[hi, k] = self.h[k]
            inside &= ~(lo | hi)
        if inside.any():
            ci = c[inside]
            dist = self.h[None, :] - np.abs(ci)
            which = np.zeros(ci.shape[0], dtype=np.int64)
            mind = dist[:, 0].copy()
            for k in range(1, ci.shape[1]):
                better = dist[:, k] < mind
                which[better] = k
                mind[better] = dist[better, k]
            rows = np.arange(ci.shape[0])
            sel = ci[rows, which]
            ci[rows, which] = np.where(sel < 0, -self.h[which], self.h[which])
            c[inside] = ci
        return self.t + c


class ComplexShape:
    """Ordered list of (shape, op) with op in {"add", "sub"} (BinaryShapes)."""

    def __init__(self):
        self.parts = []

    def add(self, shape):
        self.parts.append((shape, "add"))
        return self

    def subtract(self, shape):
        self.parts.append((shape, "sub"))
        return self

    def contains(self, p):
        exist = np.zeros(p.shape[0], dtype=bool)
        for shape, op in self.parts:
            inside = shape.contains(p)
            exist = (exist | inside) if op == "add" else (exist & ~inside)
        return exist

    def closest_point(self, p):
        best = np.full(p.shape[0], np.finfo(np.float64).max)
        out = np.zeros_like(p)
        for shape, _ in self.parts:
            found = shape.closest_point(p)
            d = np.sqrt(np.sum((p - found) ** 2, axis=1))
            take = d <= best
            best[take] = d[take]
            out[take] = found[take]
        return out


def system_bounds(lower, upper, dp):
    """SPHSystem bounds: the domain expanded by 4 dp (sph_system.cpp:39)."""
    lo = np.asarray(lower, dtype=np.float64) - dp * 4
    up = np.asarray(upper, dtype=np.float64) + dp * 4
    return lo, up


def lattice_positions(bounds_lower, bounds_upper, dp, shape):
    """Cell centres of a dp mesh over the system bounds kept if inside `shape`.

    Mesh(bounds, dp, 0): lower = bounds.lower, cells = ceil(extent / dp);
    centre = (lower + i * dp) + 0.5 * dp; loop order i, j[, k] (last fastest).
    """
    lo = np.asarray(bounds_lower, dtype=np.float64)
    up = np.asarray(bounds_upper, dtype=np.float64)
    dim = lo.size
    ncell = [int(math.ceil((up[k] + 0.0 - lo[k]) / dp)) + 1 - 1 for k in range(dim)]
    axes = [(lo[k] + np.arange(ncell[k], dtype=np.float64) * dp) + 0.5 * dp for k in range(dim)]
    # evaluate the shape slab by slab along the first axis to bound memory
    keep = []
    grids_rest = np.meshgrid(*axes[1:], indexing="ij")
    rest = np.stack([g.ravel() for g in grids_rest], axis=1)
    for x in axes[0]:
        pts = np.empty((rest.shape[0], dim), dtype=np.float64)
        pts[:, 0] = x
        pts[:, 1:] = rest
        m = shape.contains(pts)
        if m.any():
            keep.append(pts[m])
    if not keep:
        return np.zeros((0, dim), dtype=np.float64)
    return np.concatenate(keep, axis=0)


def normal_from_body_shape(shape, pos):
    """Shape::findNormalDirection for points not on the surface."""
    is_contain = shape.contains(pos)
    disp = shape.closest_point(pos) - pos
    z = np.sum(disp * disp, axis=1)
    if np.any(z <= 0.0):
        raise ValueError("normal requested on the shape surface")
    direction = disp / np.sqrt(z)[:, None]
    return np.where(is_contain[:, None], direction, -1.0 * direction)


def morton_order(pos, bounds_lower, grid_spacing, buffer_width=2):
    """Stable permutation that orders particles by the Morton code of their
    cell on the cell-linked-list mesh (the ParticleSortCK sequence,
    meshes/base_mesh.hxx:85-99, cell_linked_list.cpp:131-136).  Used to emit
    generated lattices in a cache-friendly order."""
    lower = np.asarray(bounds_lower, dtype=np.float64) - buffer_width * grid_spacing
    c = np.maximum(np.floor((pos - lower) / grid_spacing).astype(np.int64), 0)
    key = np.zeros(len(pos), dtype=np.uint64)
    for bit in range(10):
        for axis in range(pos.shape[1]):
            key |= ((c[:, axis].astype(np.uint64) >> bit) & 1) << (bit * 3 + axis)
    return np.argsort(key, kind="stable")
