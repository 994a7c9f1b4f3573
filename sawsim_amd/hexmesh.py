"""Structured Hex27 thin-slab meshes of a single-period SAW cell (host side).

Rebuilds, without Gmsh, the slabs that the reference's ``sp_hex_meshes`` scripts extrude
(``sp_hex_meshes/single.py:22-126``, ``double.py``, ``triple.py``, ``quad.py``):

* x (propagation): one wavelength w = 2 pitch, cut at the electrode edges
  w/4 -+ MR w/4 and 3w/4 -+ MR w/4 into five segments, each split into equal elements;
* y (aperture): one element across the aperture;
* z (depth): from the bottom, the PML slab (2 pitch), the backing layers (deepest first)
  and the piezo film, each in equal element layers; the two electrodes sit on the piezo
  (z from h to h + d) over segments 2 and 4.

The element node order is the reference's H27 order (``saw3d/elements.py:42``, the
``sortSF27`` permutation of the tensor-product shapes): local node ``i`` sits at natural
coordinates ``H27_NODES[i]`` with xi1, xi2, xi3 along x, y, z.

Volume tags follow the reference's physical groups: 7 piezo, 8 electrodes, 9.. the backing
layers from the top down, then the PML (``sp_hex_meshes/__init__.py:12-16``).  Boundary node
sets (``find_saw_boundary_3ds``, ``saw3d/boundary.py:37-60``) come from the geometry: L/R
(x = 0 / w), front/back (y = 0 / aperture), bottom (z = 0), zheng/fu (the surface nodes of
the positive / negative electrode volume, faces on the piezo and on front/back included).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# natural coordinates (xi1, xi2, xi3) of the 27 local nodes, reference H27 order
H27_NODES = np.array([
    (1, 1, 1), (1, -1, 1), (-1, -1, 1), (-1, 1, 1), (1, 1, -1), (1, -1, -1), (-1, -1, -1), (-1, 1, -1),
    (1, 0, 1), (0, 1, 1), (1, 1, 0), (0, -1, 1), (1, -1, 0), (-1, 0, 1), (-1, -1, 0), (-1, 1, 0),
    (1, 0, -1), (0, 1, -1), (0, -1, -1), (-1, 0, -1), (0, 0, 1), (1, 0, 0), (0, 1, 0), (0, -1, 0),
    (-1, 0, 0), (0, 0, -1), (0, 0, 0)], dtype=np.int64)

TAG_PIEZO, TAG_ELECTRODE, TAG_FIRST_LAYER = 7, 8, 9


def gmsh_segments(length, lc):
    """Elements Gmsh's 1D mesher puts on a straight curve of ``length`` with size ``lc`` at
    both ends, when the curve bounds a recombined surface: int(length / lc + 1.99) points,
    raised to the next odd count (the default Blossom recombination asks for an even number
    of segments), so the element count is always even."""
    points = max(2, int(length / lc + 1.99))
    if points % 2 == 0:
        points += 1
    return points - 1


@dataclass
class HexMesh:
    pos_m: np.ndarray               # (N, 3) metres
    hexas: np.ndarray               # (ne, 27) reference H27 local order
    hex_tag: np.ndarray             # (ne,) volume tags
    sets: dict                      # L, R, front, back, bottom, zheng, fu node arrays
    info: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return int(self.pos_m.shape[0])

    @property
    def ne(self) -> int:
        return int(self.hexas.shape[0])

    def front_section(self):
        """(quads (k, 9), tags) of the y = 0 faces: Q9 view of the slab for previews."""
        y0 = self.hexas[:, [6, 5, 1, 2, 18, 12, 11, 14, 23]]     # y = -1 face, Gmsh QUADS9-like order
        return y0, self.hex_tag.copy()


def slab_cell(pitch_um, aperture_um, electrode_um, metal_ratio, layers, pml_um, nz_pml, nz_piezo, piezo_um,
              nz_electrode, lc_um, x_counts=None):
    """Hex27 slab of one SP cell.

    ``layers`` = [(thickness_um, n_elements), ...] from just below the piezo downwards;
    ``x_counts`` = elements of the five x segments (default: Gmsh's 1D rule at ``lc_um``).
    """
    p = float(pitch_um)
    w = 2.0 * p
    mr = float(metal_ratio)
    xb = [0.0, w / 4 - w / 4 * mr, w / 4 + w / 4 * mr, w / 4 * 3 - w / 4 * mr, w / 4 * 3 + w / 4 * mr, w]
    if x_counts is None:
        x_counts = [gmsh_segments(xb[i + 1] - xb[i], lc_um) for i in range(5)]
    # x node coordinates (element edges and midpoints)
    xe = [0.0]
    for i in range(5):
        xe.extend(np.linspace(xb[i], xb[i + 1], x_counts[i] + 1)[1:].tolist())
    xe = np.array(xe)
    nx = xe.size - 1
    xn = np.empty(2 * nx + 1)
    xn[0::2] = xe
    xn[1::2] = 0.5 * (xe[:-1] + xe[1:])
    xn[-1] = w
    # electrode x node ranges (node indices, inclusive)
    seg_start = np.concatenate([[0], np.cumsum(x_counts)])
    elec = [(2 * seg_start[1], 2 * seg_start[2]), (2 * seg_start[3], 2 * seg_start[4])]
    # z slabs from the bottom: PML, layers (deepest first), piezo
    slabs = [(pml_um, nz_pml, None)]
    ntag = len(layers)
    for k in range(ntag - 1, -1, -1):
        slabs.append((layers[k][0], layers[k][1], TAG_FIRST_LAYER + k))
    slabs.append((piezo_um, nz_piezo, TAG_PIEZO))
    tag_pml = TAG_FIRST_LAYER + ntag
    ze, ztag = [0.0], []
    for thick, nz, tag in slabs:
        z0 = ze[-1]
        ze.extend((z0 + np.arange(1, nz + 1) * (thick / nz)).tolist())
        ztag.extend([tag_pml if tag is None else tag] * nz)
    h = ze[-1]
    nz_body = len(ztag)
    d = float(electrode_um)
    ze.extend((h + np.arange(1, nz_electrode + 1) * (d / nz_electrode)).tolist())
    ztag.extend([TAG_ELECTRODE] * nz_electrode)
    ze = np.array(ze)
    nz = ze.size - 1
    zn = np.empty(2 * nz + 1)
    zn[0::2] = ze
    zn[1::2] = 0.5 * (ze[:-1] + ze[1:])
    yn = np.array([0.0, 0.5 * aperture_um, float(aperture_um)])
    iz_top = 2 * nz_body
    # node ids on the (ix, iy, iz) lattice; electrode-only rows exist inside the electrode x ranges
    exists = np.zeros((2 * nx + 1, 3, 2 * nz + 1), dtype=bool)
    exists[:, :, :iz_top + 1] = True
    for a, b in elec:
        exists[a:b + 1, :, iz_top:] = True
    node_id = np.full(exists.shape, -1, dtype=np.int64)
    # numbering: z rows outermost, then y, then x
    order = np.argwhere(exists.transpose(2, 1, 0))          # (iz, iy, ix) lexicographic
    N = order.shape[0]
    node_id[order[:, 2], order[:, 1], order[:, 0]] = np.arange(N)
    pos = np.stack([xn[order[:, 2]], yn[order[:, 1]], zn[order[:, 0]]], axis=1) * 1e-6
    # elements
    ex_list, ez_list = [], []
    for ez in range(nz):
        if ez < nz_body:
            exs = range(nx)
        else:
            exs = [k for a, b in elec for k in range(a // 2, b // 2)]
        for ex in exs:
            ex_list.append(ex)
            ez_list.append(ez)
    ex_a = np.array(ex_list, dtype=np.int64)
    ez_a = np.array(ez_list, dtype=np.int64)
    ix = 2 * ex_a[:, None] + 1 + H27_NODES[None, :, 0]
    iy = 1 + H27_NODES[None, :, 1] + 0 * ex_a[:, None]
    iz = 2 * ez_a[:, None] + 1 + H27_NODES[None, :, 2]
    hexas = node_id[ix, iy, iz]
    if (hexas < 0).any():
        raise RuntimeError('hex27 slab: element references a missing node')
    hex_tag = np.array(ztag, dtype=np.int64)[ez_a]
    # boundary sets
    lat = np.argwhere(exists)                                # (ix, iy, iz)
    ids = node_id[lat[:, 0], lat[:, 1], lat[:, 2]]

    def where(mask):
        return np.sort(ids[mask])
    sets = dict(L=where(lat[:, 0] == 0), R=where(lat[:, 0] == 2 * nx), front=where(lat[:, 1] == 0),
                back=where(lat[:, 1] == 2), bottom=where(lat[:, 2] == 0))
    for name, (a, b) in zip(('zheng', 'fu'), elec):
        inside = (lat[:, 0] >= a) & (lat[:, 0] <= b) & (lat[:, 2] >= iz_top)
        surface = ((lat[:, 0] == a) | (lat[:, 0] == b) | (lat[:, 1] == 0) | (lat[:, 1] == 2)
                   | (lat[:, 2] == iz_top) | (lat[:, 2] == 2 * nz))
        sets[name] = where(inside & surface)
    tag_map = {TAG_PIEZO: 0, TAG_ELECTRODE: 1}
    for k in range(ntag):
        tag_map[TAG_FIRST_LAYER + k] = k + 2
    tag_map[tag_pml] = ntag + 1 if ntag else 0
    info = dict(period_m=w * 1e-6, aperture_m=float(aperture_um) * 1e-6, pml_thickness_m=float(pml_um) * 1e-6,
                tag_map=tag_map, tag_pml=tag_pml, x_counts=list(map(int, x_counts)),
                z_counts=[int(s[1]) for s in slabs] + [int(nz_electrode)], top_m=h * 1e-6)
    return HexMesh(pos_m=pos, hexas=hexas, hex_tag=hex_tag, sets=sets, info=info)


def _nn(k, h, h_ref, lc_ref, lc):
    """Layer count rule of the reference slab scripts: ceil(k h / h_ref * lc_ref / lc)."""
    return max(1, math.ceil(k * h / h_ref * lc_ref / lc - 1e-12))


def template_slab(name, pitch_um, aperture_um, electrode_um, substrate_um, metal_ratio, mesh_um, layers_um=(),
                  x_counts=None):
    """The reference's sp_2p5d_{single,double,triple,quad}_layer slab (Hex27, one element in y).

    Layer counts follow ``sp_hex_meshes/{single,double,triple,quad}.py`` (``NN`` rules);
    the PML is 2 pitch thick.
    """
    p, lc, d = pitch_um, mesh_um, electrode_um
    h_pml = 2 * p
    if name == 'single':
        nn = _nn(1, d, 0.17, 0.217, lc)
        n1 = _nn(30, substrate_um, 6.51, 0.217, lc)
        n_pml = _nn(10, h_pml, 2.17, 0.217, lc)
        layers = []
    elif name == 'double':
        nn = _nn(1, d, 0.17, 0.217, lc)
        n1 = _nn(3, substrate_um, 0.6, 0.217, lc)
        layers = [(layers_um[0], _nn(30, layers_um[0], 6.51, 0.217, lc))]
        n_pml = _nn(10, h_pml, 2.17, 0.217, lc)
    elif name == 'triple':
        nn = _nn(1, d, 0.17, 0.217, lc)
        n1 = _nn(3, substrate_um, 0.6, 0.217, lc)
        layers = [(layers_um[0], _nn(3, layers_um[0], 0.5, 0.217, lc)),
                  (layers_um[1], _nn(30, layers_um[1], 6.51, 0.217, lc))]
        n_pml = _nn(11, h_pml, 2.17, 0.217, lc)
    elif name == 'quad':
        nn = _nn(1, d, 0.17, 0.2, lc)
        n1 = _nn(3, substrate_um, 0.6, 0.2, lc)
        layers = [(layers_um[0], _nn(3, layers_um[0], 0.5, 0.2, lc)),
                  (layers_um[1], _nn(5, layers_um[1], 1.0, 0.2, lc)),
                  (layers_um[2], _nn(30, layers_um[2], 6.0, 0.2, lc))]
        n_pml = _nn(10, h_pml, 2.0, 0.2, lc)
    else:
        raise ValueError(f'unknown slab template {name!r}')
    return slab_cell(p, aperture_um, d, metal_ratio, layers, h_pml, n_pml, n1, substrate_um, nn, lc,
                     x_counts=x_counts)
