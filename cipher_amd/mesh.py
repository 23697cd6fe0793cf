"""Host-side leaf forest and its flattening into the right-hand side's lists.

CIPHER keeps its mesh in a p4est brick forest converted to DMPlex and, after
every adapt, flattens it into owned cells, interior faces with two supports
and no tree children (hanging child faces kept, parent faces dropped) and a
per-cell size h = pow(volume, 1/dim) (src/cipher.c:1187-1231).  This module
builds the equivalent on the host with numpy:

* ``Forest``: a 2:1-balanced (corner-balanced) brick of quadtrees/octrees in
  finest-voxel coordinates, leaves in Morton order of their anchors;
* ``adapt_to_labels``: top-down refinement at label interfaces between the
  minimum and maximum levels, followed by ripple 2:1 balancing;
* ``Forest.flatten``: the face list (L = low side, R = high side of the face),
  boundary faces with the label ids of src/cipher.c:1214-1220, and h;
* ``adapt`` / ``transfer``: the time loop's remesh (src/cipher.c:1607-1700):
  refine and coarsen by label within the level limits, 2:1 balance, and the
  state carried over (copied into refined leaves, averaged into coarsened
  ones).

Faces are ordered by (L, direction, R); cell ordering and face ordering are
this module's own (DMPlex numbering is not reproduced).
"""
import math

import numpy as np
from scipy import ndimage


def morton_keys(coords, bits):
    """Interleave the bits of integer coords (N x dim) into Morton keys."""
    coords = np.asarray(coords, dtype=np.uint64)
    n, dim = coords.shape
    key = np.zeros(n, dtype=np.uint64)
    for b in range(bits):
        for d in range(dim):
            key |= ((coords[:, d] >> np.uint64(b)) & np.uint64(1)) << np.uint64(b * dim + d)
    return key


class Forest:
    """Leaves of a brick forest: level (per tree) and anchor (finest voxels)."""

    def __init__(self, res, blocks, size, periodic, levels, anchors):
        self.res = np.asarray(res, dtype=np.int64)
        self.dim = len(res)
        self.blocks = np.asarray(blocks, dtype=np.int64)
        self.size = np.asarray(size, dtype=np.float64)
        self.periodic = list(periodic)
        self.lmax = int(round(math.log2(int(self.res[0] // self.blocks[0]))))
        for d in range(self.dim):
            if int(self.blocks[d]) << self.lmax != int(self.res[d]):
                raise ValueError("grid must be blocks * 2^level in every dimension")
        lv = np.asarray(levels, dtype=np.int64)
        an = np.asarray(anchors, dtype=np.int64).reshape(-1, self.dim)
        bits = max(1, int(math.ceil(math.log2(max(int(self.res.max()), 2)))))
        order = np.argsort(morton_keys(an, bits), kind="stable")
        self.level = lv[order]
        self.anchor = an[order]

    @property
    def ncells(self):
        return int(self.level.size)

    def cell_extent(self):
        """Cell edge length in voxels."""
        return np.left_shift(np.int64(1), self.lmax - self.level)

    def id_map(self):
        """Leaf index covering every finest voxel (array indexed [x,y(,z)])."""
        m = np.empty(tuple(self.res), dtype=np.int64)
        ext = self.cell_extent()
        for l in np.unique(self.level):
            sel = np.flatnonzero(self.level == l)
            s = int(ext[sel[0]])
            blk = tuple(int(r) // s for r in self.res)
            tmp = np.empty(blk, dtype=np.int64)
            tmp[tuple((self.anchor[sel] // s).T)] = sel
            # expand every block id to its s^dim voxels (only the voxels of these leaves are used)
            mask = np.zeros(blk, dtype=bool)
            mask[tuple((self.anchor[sel] // s).T)] = True
            full = tmp
            fmask = mask
            for d in range(self.dim):
                full = np.repeat(full, s, axis=d)
                fmask = np.repeat(fmask, s, axis=d)
            m[fmask] = full[fmask]
        return m

    def level_map(self):
        return self.level[self.id_map()]

    def centroid_voxels(self):
        ext = self.cell_extent()
        return self.anchor + (ext // 2)[:, None]

    def cellgeom(self):
        """h = pow(volume, 1/dim) (src/cipher.c:1229) per cell."""
        dx = self.size / self.res
        h = np.empty(self.ncells, dtype=np.float64)
        ext = self.cell_extent()
        for l in np.unique(self.level):
            s = float(1 << int(self.lmax - l))
            vol = 1.0
            for d in range(self.dim):
                vol *= dx[d] * s
            h[self.level == l] = math.pow(vol, 1.0 / self.dim)
        return h

    def flatten(self):
        """Interior faces (2*nf int32, L then R), boundary faces and h."""
        dim = self.dim
        idm = self.id_map()
        ext = self.cell_extent()
        faces, dirs, bcell, bid = [], [], [], []
        idx = np.arange(self.ncells, dtype=np.int64)
        for d in range(dim):
            # + direction: neighbour voxel at anchor + ext along d
            pos = self.anchor.copy()
            pos[:, d] += ext
            at_hi = pos[:, d] >= self.res[d]
            if self.periodic[d]:
                pos[:, d] %= self.res[d]
                inside = np.ones(self.ncells, dtype=bool)
            else:
                inside = ~at_hi
                bcell.append(idx[at_hi]); bid.append(np.full(int(at_hi.sum()), 2 * d + 2))
            nb = np.full(self.ncells, -1, dtype=np.int64)
            nb[inside] = idm[tuple(pos[inside].T)]
            ok = inside.copy()
            ok[inside] &= self.level[nb[inside]] <= self.level[inside]
            faces.append(np.stack([idx[ok], nb[ok]], axis=1)); dirs.append(np.full(int(ok.sum()), d))
            # - direction, coarser neighbour only (hanging face seen from the fine cell)
            pos = self.anchor.copy()
            pos[:, d] -= 1
            at_lo = pos[:, d] < 0
            if self.periodic[d]:
                pos[:, d] %= self.res[d]
                inside = np.ones(self.ncells, dtype=bool)
            else:
                inside = ~at_lo
                bcell.append(idx[at_lo]); bid.append(np.full(int(at_lo.sum()), 2 * d + 1))
            nb = np.full(self.ncells, -1, dtype=np.int64)
            nb[inside] = idm[tuple(pos[inside].T)]
            ok = inside.copy()
            ok[inside] &= self.level[nb[inside]] < self.level[inside]
            faces.append(np.stack([nb[ok], idx[ok]], axis=1)); dirs.append(np.full(int(ok.sum()), d))
        F = np.concatenate(faces, axis=0)
        Dd = np.concatenate(dirs)
        order = np.lexsort((F[:, 1], Dd, F[:, 0]))
        F = F[order]
        if bcell:
            bc = np.concatenate(bcell)
            bi = np.concatenate(bid)
            o = np.lexsort((bc, bi))
            bc, bi = bc[o], bi[o]
        else:
            bc = np.zeros(0, dtype=np.int64)
            bi = np.zeros(0, dtype=np.int64)
        period = [float(self.size[d]) if self.periodic[d] else 0.0 for d in range(dim)]
        return FlatMesh(F.astype(np.int32), self.cellgeom(), bc.astype(np.int32), bi.astype(np.int32),
                        dim=self.dim, centroid=self.centroids(), period=period)

    def centroids(self):
        """Cell centres in physical units, (ncells, dim)."""
        dx = self.size / self.res
        ext = self.cell_extent().astype(np.float64)
        return (self.anchor.astype(np.float64) + 0.5 * ext[:, None]) * dx[None, :]


class FlatMesh:
    """Flattened mesh arrays in the layout of cipher_mesh."""

    def __init__(self, faces, cellgeom, bface_cell, bface_id, dim, centroid=None, period=None):
        self.faces = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 2)
        self.cellgeom = np.ascontiguousarray(cellgeom, dtype=np.float64)
        self.bface_cell = np.ascontiguousarray(bface_cell, dtype=np.int32)
        self.bface_id = np.ascontiguousarray(bface_id, dtype=np.int32)
        self.dim = dim
        self.centroid = None if centroid is None else np.ascontiguousarray(centroid, dtype=np.float64).reshape(-1)
        self.period = list(period or []) + [0.0] * (3 - len(period or []))

    @property
    def ncells(self):
        return int(self.cellgeom.size)

    @property
    def nfaces(self):
        return int(self.faces.shape[0])

    def struct(self):
        import ctypes as C
        from . import abi
        m = abi.Mesh()
        m.face_cells = self.faces.ctypes.data_as(C.POINTER(C.c_int32))
        m.cellgeom = self.cellgeom.ctypes.data_as(C.POINTER(C.c_double))
        m.bface_cell = self.bface_cell.ctypes.data_as(C.POINTER(C.c_int32))
        m.bface_id = self.bface_id.ctypes.data_as(C.POINTER(C.c_int32))
        m.ncells, m.nfaces, m.nbfaces = self.ncells, self.nfaces, int(self.bface_cell.size)
        if self.centroid is not None:
            m.centroid = self.centroid.ctypes.data_as(C.POINTER(C.c_double))
        for d in range(3):
            m.period[d] = self.period[d]
        m._keep = self
        return m


def uniform_forest(res, blocks, size, periodic, level=None):
    res = np.asarray(res, dtype=np.int64)
    blocks = np.asarray(blocks, dtype=np.int64)
    lmax = int(round(math.log2(int(res[0] // blocks[0]))))
    level = lmax if level is None else level
    s = 1 << (lmax - level)
    grids = np.meshgrid(*[np.arange(0, int(r), s) for r in res], indexing="ij")
    anchors = np.stack([g.ravel() for g in grids], axis=1)
    return Forest(res, blocks, size, periodic, np.full(anchors.shape[0], level), anchors)


def _filter_mode(periodic):
    return tuple("wrap" if p else "nearest" for p in periodic)


def _block_reduce(a, s, fn):
    if s == 1:
        return a
    dim = a.ndim
    shp = []
    for r in a.shape:
        shp += [r // s, s]
    return fn(a.reshape(shp), axis=tuple(range(1, 2 * dim, 2)))


def balance(forest):
    """Ripple 2:1 balance across faces, edges and corners: refine every leaf
    that has a neighbour more than one level finer."""
    f = forest
    dim = f.dim
    mode = _filter_mode(f.periodic)
    while True:
        lm = f.level_map()
        mx = ndimage.maximum_filter(lm, size=3, mode=mode)
        ext = f.cell_extent()
        refine = np.zeros(f.ncells, dtype=bool)
        for l in np.unique(f.level):
            sel = np.flatnonzero(f.level == l)
            s = int(ext[sel[0]])
            pooled = _block_reduce(mx, s, np.max)
            nbmax = pooled[tuple((f.anchor[sel] // s).T)]
            refine[sel] = nbmax > l + 1
        if not refine.any():
            return f
        f = _refine(f, refine)


def _refine(f, mask):
    dim = f.dim
    ext = f.cell_extent()
    keep_l = f.level[~mask]
    keep_a = f.anchor[~mask]
    lv = f.level[mask] + 1
    an = f.anchor[mask]
    half = ext[mask] // 2
    offs = np.array(np.meshgrid(*[[0, 1]] * dim, indexing="ij")).reshape(dim, -1).T
    ch_a = (an[:, None, :] + offs[None, :, :] * half[:, None, None]).reshape(-1, dim)
    ch_l = np.repeat(lv, offs.shape[0])
    return Forest(f.res, f.blocks, f.size, f.periodic,
                  np.concatenate([keep_l, ch_l]), np.concatenate([keep_a, ch_a]))


def adapt_to_labels(labels, blocks, size, periodic, minlevel, maxlevel):
    """Top-down refinement: a cell is refined below maxlevel when the labels
    in its region, widened by one finest voxel, are not all equal, and always
    below minlevel; then 2:1 balance.  labels is indexed [x,y(,z)]."""
    labels = np.asarray(labels)
    res = np.asarray(labels.shape, dtype=np.int64)
    dim = labels.ndim
    blocks = np.asarray(blocks, dtype=np.int64)
    lmax = int(round(math.log2(int(res[0] // blocks[0]))))
    maxlevel = min(maxlevel, lmax)
    mode = _filter_mode(periodic)
    emin = ndimage.minimum_filter(labels, size=3, mode=mode)
    emax = ndimage.maximum_filter(labels, size=3, mode=mode)
    leaves_l, leaves_a = [], []
    s0 = 1 << lmax
    cur = np.stack([g.ravel() for g in np.meshgrid(*[np.arange(0, int(r), s0) for r in res],
                                                   indexing="ij")], axis=1)
    offs = np.array(np.meshgrid(*[[0, 1]] * dim, indexing="ij")).reshape(dim, -1).T
    for l in range(0, maxlevel + 1):
        s = 1 << (lmax - l)
        if l < maxlevel:
            bmin = _block_reduce(emin, s, np.min)[tuple((cur // s).T)]
            bmax = _block_reduce(emax, s, np.max)[tuple((cur // s).T)]
            ref = (bmin != bmax) | (l < minlevel)
        else:
            ref = np.zeros(cur.shape[0], dtype=bool)
        leaves_l.append(np.full(int((~ref).sum()), l))
        leaves_a.append(cur[~ref])
        half = s // 2
        cur = (cur[ref][:, None, :] + offs[None] * half).reshape(-1, dim)
        if cur.size == 0:
            break
    f = Forest(res, blocks, size, periodic, np.concatenate(leaves_l), np.concatenate(leaves_a))
    return balance(f)


def adapt(forest, refine, minlevel, maxlevel):
    """DMAdaptLabel on the forest (src/cipher.c:1616-1668, level limits from
    DMForestSetMinimum/MaximumRefinement, src/cipher.c:1153-1154): a leaf
    labelled for refinement (refine[i] True) is split once if it is below
    maxlevel; every other leaf carries the label's default, coarsen, and a
    family of 2^dim sibling leaves that all carry it is merged into its
    parent if that stays at or above minlevel; then 2:1 balance.  One level
    per call either way, as p4est's non-recursive refine and coarsen."""
    f = forest
    dim = f.dim
    refine = np.asarray(refine, dtype=bool)
    if refine.shape != (f.ncells,):
        raise ValueError("one refinement flag per leaf")
    ext = f.cell_extent()
    # families: leaves of one level sharing the parent anchor
    pext = 2 * ext
    parent = (f.anchor // pext[:, None]) * pext[:, None]
    bits = max(1, int(math.ceil(math.log2(max(int(f.res.max()), 2)))))
    key = morton_keys(parent, bits).astype(np.int64) * 64 + f.level
    coarsen = (f.level > minlevel) & ~refine
    uk, inv, cnt = np.unique(key, return_inverse=True, return_counts=True)
    ok = np.ones(uk.size, dtype=bool)
    np.logical_and.at(ok, inv, coarsen)
    merge = ok[inv] & (cnt[inv] == 1 << dim)
    split = refine & (f.level < min(maxlevel, f.lmax))
    keep = ~merge & ~split
    lv = [f.level[keep]]
    an = [f.anchor[keep]]
    if split.any():
        offs = np.array(np.meshgrid(*[[0, 1]] * dim, indexing="ij")).reshape(dim, -1).T
        half = ext[split] // 2
        an.append((f.anchor[split][:, None, :] + offs[None, :, :] * half[:, None, None]).reshape(-1, dim))
        lv.append(np.repeat(f.level[split] + 1, offs.shape[0]))
    if merge.any():
        first = merge & np.all(f.anchor == parent, axis=1)      # one entry per merged family
        an.append(parent[first])
        lv.append(f.level[first] - 1)
    g = Forest(f.res, f.blocks, f.size, f.periodic, np.concatenate(lv), np.concatenate(an))
    return balance(g)


def transfer(old, new, X):
    """DMForestTransferVec (src/cipher.c:1675) of cell records X (old.ncells
    rows): a new leaf inside an old leaf takes its record (refinement
    copies); a new leaf covering old leaves takes their mean, summed over
    the children in (x, y, z) index order, z fastest (coarsening averages
    equal volumes)."""
    X = np.asarray(X)
    if X.shape[0] != old.ncells:
        raise ValueError("X must have one row per old leaf")
    dim = new.dim
    idm = old.id_map()
    cover = idm[tuple(new.anchor.T)]
    out = np.empty((new.ncells,) + X.shape[1:], dtype=X.dtype)
    same = old.level[cover] <= new.level
    out[same] = X[cover[same]]
    rest = np.flatnonzero(~same)
    if rest.size:
        offs = np.array(np.meshgrid(*[[0, 1]] * dim, indexing="ij")).reshape(dim, -1).T
        half = new.cell_extent()[rest] // 2
        ch = new.anchor[rest][:, None, :] + offs[None, :, :] * half[:, None, None]
        ids = idm[tuple(ch.reshape(-1, dim).T)].reshape(rest.size, -1)
        if np.any(old.level[ids] != (new.level[rest] + 1)[:, None]):
            raise ValueError("a coarsened leaf must cover one family of old leaves")
        acc = X[ids[:, 0]].copy()
        for k in range(1, ids.shape[1]):
            acc += X[ids[:, k]]
        out[rest] = acc / ids.shape[1]
    return out


def cell_labels(forest, labels):
    """Label of each leaf at its centroid voxel (src/cipher.c:1297-1308)."""
    cv = forest.centroid_voxels()
    return np.asarray(labels)[tuple(cv.T)].astype(np.int32)
