"""Structured Q9 unit-cell meshes for periodic SAW cells (no Gmsh).

The reference builds its structured templates with Gmsh transfinite surfaces
(``sp_meshes/stack.py:87-177``, ``sp_meshes/double.py:21-176``,
``sp_meshes/quad.py``).  For straight-sided transfinite blocks that mesh is a
tensor product: each vertical strip of the cell is split into ``cols`` equal
elements, each horizontal band into ``rows`` equal elements, and the second-order
nodes sit at the edge and face midpoints.  This module writes that tensor
product directly.

Output follows the reference ``piezo_fem.Model`` inputs (``piezo_fem.py:161-186``):

* ``pos_m``  (N, 2) node coordinates in metres,
* ``quads``  (ne, 9) node ids, Gmsh QUADS9 order (corners counter-clockwise from
  the lower-left, then the mid-edge nodes of edges 0-1, 1-2, 2-3, 3-0, then the
  centre), matching ``_q9_shapes`` (``piezo_fem.py:56-103``),
* ``elem_tag`` (ne,) physical surface tags (15 electrode, 16 piezo, 22.. layers,
  then the PML tag),
* ``lines``  (nl, 4) 3-node boundary segments ``[n1, n2, n_mid, tag]`` with
  1-based node ids, read by ``nodes_with_tag`` (``piezo_fem.py:174-177``).
  Line tags: 17 positive electrode, 18 negative electrode, 19 left, 20 right,
  21 bottom (``sp_meshes/stack.py:166-171``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

TAG_ELECTRODE = 15
TAG_PIEZO = 16
LINE_POS, LINE_NEG, LINE_LEFT, LINE_RIGHT, LINE_BOTTOM = 17, 18, 19, 20, 21


@dataclass
class CellMesh:
    """A meshed periodic unit cell (the reference's ``Model`` plus its line tags)."""

    pos_m: np.ndarray
    quads: np.ndarray
    elem_tag: np.ndarray
    lines: np.ndarray
    info: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return int(self.pos_m.shape[0])

    @property
    def ne(self) -> int:
        return int(self.quads.shape[0])

    def nodes_with_tag(self, line_tag: int) -> np.ndarray:
        """Sorted unique 0-based nodes on boundary segments with ``line_tag``."""
        sel = self.lines[self.lines[:, 3] == line_tag][:, :3].astype(np.int64) - 1
        return np.unique(sel)


def _strip_coords(cuts, counts):
    """Node coordinates (with Q9 midpoints) along one axis of a tensor mesh."""
    out = [float(cuts[0])]
    for i, n in enumerate(counts):
        a, b = float(cuts[i]), float(cuts[i + 1])
        for k in range(1, 2 * n + 1):
            out.append(a + (b - a) * k / (2 * n))
    return np.asarray(out)


def structured_cell(xs, cols, levels, rows, region, electrodes, info=None) -> CellMesh:
    """Tensor-product Q9 mesh of a strip/band block layout.

    Parameters
    ----------
    xs : x cut positions (µm), len = nstrips + 1; cell spans [xs[0], xs[-1]].
    cols : elements per strip.
    levels : y cut positions (µm), bottom to top, len = nbands + 1.
    rows : elements per band.
    region : callable (band, strip) -> surface tag or None (no elements there).
    electrodes : dict line_tag -> (band, first_strip, last_strip): the electrode
        occupies strips first..last of that band; its boundary (bottom, two sides,
        top) carries the tag, as in ``sp_meshes/stack.py:166-168``.
    """
    xg = _strip_coords(xs, cols)
    yg = _strip_coords(levels, rows)
    col0 = np.concatenate([[0], np.cumsum(cols)])
    row0 = np.concatenate([[0], np.cumsum(rows)])
    NXg, NYg = xg.size, yg.size

    elems, tags = [], []
    for b in range(len(rows)):
        for s in range(len(cols)):
            t = region(b, s)
            if t is None:
                continue
            for r in range(rows[b]):
                iy = 2 * (row0[b] + r)
                for c in range(cols[s]):
                    ix = 2 * (col0[s] + c)
                    elems.append((ix, iy))
                    tags.append(t)
    elems = np.asarray(elems, dtype=np.int64)
    ix, iy = elems[:, 0], elems[:, 1]
    # grid offsets of the nine local nodes, QUADS9 order
    off = np.array([(0, 0), (2, 0), (2, 2), (0, 2), (1, 0), (2, 1), (1, 2), (0, 1), (1, 1)])
    gx = ix[:, None] + off[None, :, 0]
    gy = iy[:, None] + off[None, :, 1]
    gid = gy * NXg + gx                                          # grid id, row-major from bottom
    used = np.unique(gid)
    remap = np.full(NXg * NYg, -1, dtype=np.int64)
    remap[used] = np.arange(used.size)
    quads = remap[gid]
    pos_um = np.column_stack([xg[used % NXg], yg[used // NXg]])

    def seg(gx0, gy0, gx1, gy1):
        a = remap[gy0 * NXg + gx0]
        b = remap[gy1 * NXg + gx1]
        m = remap[((gy0 + gy1) // 2) * NXg + (gx0 + gx1) // 2]
        return a, b, m

    lines = []
    xmax = 2 * col0[-1]
    for (ex, ey) in elems:
        if ex == 0:
            lines.append((*seg(ex, ey, ex, ey + 2), LINE_LEFT))
        if ex + 2 == xmax:
            lines.append((*seg(ex + 2, ey, ex + 2, ey + 2), LINE_RIGHT))
        if ey == 0:
            lines.append((*seg(ex, ey, ex + 2, ey), LINE_BOTTOM))
    for tag, (band, s0, s1) in electrodes.items():
        ylo, yhi = 2 * row0[band], 2 * row0[band + 1]
        xlo, xhi = 2 * col0[s0], 2 * col0[s1 + 1]
        for gx0 in range(xlo, xhi, 2):                            # bottom and top
            lines.append((*seg(gx0, ylo, gx0 + 2, ylo), tag))
            lines.append((*seg(gx0, yhi, gx0 + 2, yhi), tag))
        for gy0 in range(ylo, yhi, 2):                            # both sides
            lines.append((*seg(xlo, gy0, xlo, gy0 + 2), tag))
            lines.append((*seg(xhi, gy0, xhi, gy0 + 2), tag))
    lines = np.asarray(lines, dtype=np.int64)
    lines[:, :3] += 1                                             # 1-based like Gmsh
    return CellMesh(pos_m=pos_um * 1e-6, quads=quads, elem_tag=np.asarray(tags, dtype=np.int64),
                    lines=lines, info=dict(info or {}, cols=list(cols), rows=list(rows)))


def _electrode_xs(p, mr):
    """Nine x cuts at the electrode edges and centres (``sp_meshes/stack.py:122``)."""
    return [0.0, p * (1 - mr) / 2, p / 2, p * (1 + mr) / 2, p, p * (3 - mr) / 2, 1.5 * p,
            p * (3 + mr) / 2, 2 * p]


def stack_cell(pitch_um, metal_ratio, electrode_um, piezo_um, backings=(), coatings=(),
               mesh_um=0.5, pml_pitches=2.0) -> CellMesh:
    """Generic layered cell with the ``sp_stack`` meshing rules.

    Follows ``sp_meshes/stack.py:87-123``: 8 strips between the electrode cuts with
    ``ceil(width/h)`` elements each; below the surface every band gets
    ``ceil(height/h)`` rows within ``2·pitch`` of the surface and ``h_deep`` rows
    deeper (bands crossing that depth are split); the bottom PML
    (``pml_pitches·pitch``) gets ``max(8, ceil(pml/h_deep))`` rows; above the
    surface the electrode band and the coatings (the first coating is measured from
    the piezo surface and embeds the electrodes).

    Surface tags: 16 piezo, 15 electrode, 22.. backings (top to bottom), then
    coatings (inner to outer), then the PML.
    """
    p, h, d = float(pitch_um), float(mesh_um), float(electrode_um)
    h_deep = max(h, min(4 * h, p / 2))
    near = 2 * p
    nb, nc = len(backings), len(coatings)
    tag_backing = [22 + i for i in range(nb)]
    tag_coating = [22 + nb + j for j in range(nc)]
    tag_pml = 22 + nb + nc
    bands, z = [(TAG_PIEZO, 0.0, float(piezo_um))], float(piezo_um)
    for tag, t in zip(tag_backing, backings):
        bands.append((tag, z, z + float(t)))
        z += float(t)
    split = []
    for tag, top, bot in bands:
        if top < near < bot - 1e-12:
            split += [(tag, top, near), (tag, near, bot)]
        else:
            split.append((tag, top, bot))
    pml_h = pml_pitches * p
    below = [(tag, -bot, -top, max(1, math.ceil((bot - top) / (h if bot <= near + 1e-12 else h_deep) - 1e-9)))
             for tag, top, bot in split]
    below.append((tag_pml, -(z + pml_h), -z, max(8, math.ceil(pml_h / h_deep - 1e-9))))
    below.sort(key=lambda b: b[1])
    above = [('E', 0.0, d, max(1, math.ceil(d / h - 1e-9)))]
    if nc:
        top = float(coatings[0])
        above.append((tag_coating[0], d, top, max(1, math.ceil((top - d) / h - 1e-9))))
        for tag, t in zip(tag_coating[1:], coatings[1:]):
            above.append((tag, top, top + float(t), max(1, math.ceil(float(t) / h - 1e-9))))
            top += float(t)
    levels = [b[1] for b in below] + [0.0] + [b[2] for b in above]
    rows = [b[3] for b in below] + [b[3] for b in above]
    band_tag = [b[0] for b in below] + [b[0] for b in above]
    ebands = len(below)
    xs = _electrode_xs(p, metal_ratio)
    cols = [max(1, math.ceil((xs[i + 1] - xs[i]) / h - 1e-9)) for i in range(8)]
    estrip = {1, 2, 5, 6}

    def region(b, s):
        if b == ebands:
            if s in estrip:
                return TAG_ELECTRODE
            return tag_coating[0] if nc else None
        return band_tag[b]

    info = dict(template='sp_stack', tag_backing=tag_backing, tag_coating=tag_coating, tag_pml=tag_pml,
                pml_thickness_m=pml_h * 1e-6, period_m=2 * p * 1e-6, h_near=h, h_deep=h_deep)
    return structured_cell(xs, cols, levels, rows, region,
                           {LINE_POS: (ebands, 1, 2), LINE_NEG: (ebands, 5, 6)}, info)


def layered_template_cell(pitch_um, electrode_um, metal_ratio, band_heights, band_rows,
                          electrode_rows, strip_cols, tags) -> CellMesh:
    """Fixed-row layered templates (``sp_meshes/double.py``, ``sp_meshes/quad.py``).

    ``band_heights``/``band_rows``/``tags`` list the substrate bands bottom to top
    (PML first, piezo last); the electrode teeth (one row block of
    ``electrode_rows``) sit on strips 1,2 and 5,6 above the surface.
    """
    p = float(pitch_um)
    xs = _electrode_xs(p, metal_ratio) if metal_ratio != 0.5 else [i * 2 * p / 8 for i in range(9)]
    levels = [0.0]
    for hgt in reversed(band_heights):
        levels.insert(0, levels[0] - float(hgt))
    levels.append(float(electrode_um))
    rows = list(band_rows) + [electrode_rows]
    nb = len(band_heights)

    def region(b, s):
        if b == nb:
            return TAG_ELECTRODE if s in (1, 2, 5, 6) else None
        return tags[b]

    return structured_cell(xs, [strip_cols] * 8, levels, rows, region,
                           {LINE_POS: (nb, 1, 2), LINE_NEG: (nb, 5, 6)},
                           dict(period_m=2 * p * 1e-6))


def double_layer_cell(pitch=1.085, electrode_um=0.17, metal_ratio=0.5, h_lt=0.6, h_si=8.33,
                      mesh_um=0.5) -> CellMesh:
    """``sp_double_layer`` mesh (piezo / Si / PML, ``sp_meshes/double.py:21-176``).

    Row counts follow ``double.py:48-52``; tags 16 piezo, 22 Si, 23 PML, 15 Al.
    """
    w = 2 * pitch
    nx = max(1, math.ceil(.5 / mesh_um - 1e-12))
    ny_lt = max(1, math.ceil(3 * h_lt / 0.6 * .5 / mesh_um - 1e-12))
    ny_si = max(1, math.ceil(31 * h_si / 8.33 * .5 / mesh_um - 1e-12))
    ny_pml = max(1, math.ceil(8 * w / 2.17 * .5 / mesh_um - 1e-12))
    ny_al = max(1, math.ceil(1 * electrode_um / 0.17 * .5 / mesh_um - 1e-12))
    m = layered_template_cell(pitch, electrode_um, metal_ratio, [w, h_si, h_lt], [ny_pml, ny_si, ny_lt],
                              ny_al, nx, [23, 22, 16])
    m.info.update(template='sp_double_layer', tag_pml=23, pml_thickness_m=w * 1e-6,
                  tag_map={16: 0, 15: 1, 22: 2, 23: 2})
    return m


def quad_layer_cell(pitch=1.085, metal_ratio=0.5, electrode_um=0.17, h_lt=0.6, h_sio2=0.5,
                    h_polysi=1.0, h_si=7.83, mesh_um=0.5) -> CellMesh:
    """``sp_quad_layer`` mesh (piezo / SiO2 / poly-Si / Si / PML, ``sp_meshes/quad.py``).

    Tags 16 piezo, 22 SiO2, 23 poly-Si, 24 Si, 25 PML, 15 Al; PML = 2·pitch.
    """
    w = 2 * pitch
    nx = max(1, math.ceil(.5 / mesh_um - 1e-12))

    def ny(rows, height, default):
        return max(1, math.ceil(rows * height / default * .5 / mesh_um - 1e-12))

    m = layered_template_cell(pitch, electrode_um, metal_ratio,
                              [w, h_si, h_polysi, h_sio2, h_lt],
                              [ny(8, w, 2.17), ny(29, h_si, 7.83), ny(4, h_polysi, 1.), ny(2, h_sio2, .5),
                               ny(3, h_lt, .6)],
                              ny(1, electrode_um, .17), nx, [25, 24, 23, 22, 16])
    m.info.update(template='sp_quad_layer', tag_pml=25, pml_thickness_m=w * 1e-6,
                  tag_map={16: 0, 15: 1, 22: 2, 23: 3, 24: 4, 25: 4})
    return m


def stack_tag_map(mesh: CellMesh) -> dict:
    """Surface tag -> material index for ``stack_cell`` (``sp_meshes/__init__.py:31-34``).

    Index 0 piezo, 1 electrode, 2.. backings then coatings; the PML continues the
    lowest backing layer, or the piezo when there is none.
    """
    tb, tc, tp = mesh.info['tag_backing'], mesh.info['tag_coating'], mesh.info['tag_pml']
    tags = tb + tc
    mapping = {TAG_PIEZO: 0, TAG_ELECTRODE: 1, tp: (1 + len(tb)) if tb else 0}
    mapping.update({tag: i + 2 for i, tag in enumerate(tags)})
    return mapping
