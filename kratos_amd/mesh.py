"""Uniform block meshes and the device context.

``MeshSpec`` reproduces the reference's Cartesian block generator for a
same-level mesh: block order, per-block geometry (meshgen.cpp:286-321,
``loc(1) - loc(0)`` in the same float arithmetic) and the 26-neighbour table
with periodic wrap (tree.cpp:77-84, 177-197).  ``Context`` owns one
``kr_ctx`` of the C-ABI and moves global numpy fields in and out of it
through the per-block AoS layout of the reference's host arrays.

Global field layout used by the Python side:
  u  : (5, Nz, Ny, Nx) conserved rho, E, m0, m1, m2
  bx : (Nz, Ny, Nx+1), by : (Nz, Ny+1, Nx), bz : (Nz+1, Ny, Nx) face fields
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _capi


def bf_faces(n):
    """Face indices per axis of the face-field array: n + 1 on an active
    axis, 1 on a thin (one-cell) axis (mhd/types/fld_data.cpp:18-26)."""
    return tuple(int(v) + (1 if v > 1 else 0) for v in n)


def bf_host_dims(n):
    """(N0, N1, N2) of a block's host face-field array: the faces plus two
    ghost layers on each side of every axis."""
    return tuple(f + 4 for f in bf_faces(n))


def bf_host_size(n):
    N = bf_host_dims(n)
    return N[0] * N[1] * N[2] * 3


def _divb_cells(bfs, dx):
    """|div B| per cell from face fields bx (nz, ny, nx+1), by (nz, ny+1,
    nx), bz (nz+1, ny, nx); a thin axis (one face, both sides of the cell)
    contributes nothing."""
    d = 0.0
    for a, f in enumerate(bfs):
        ax = 2 - a                                   # arrays are [z, y, x]
        if f.shape[ax] < 2:
            continue
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        cell = [min(g.shape[k] for g in bfs) for k in range(3)]
        cut = tuple(slice(0, cell[k]) for k in range(3))
        d = d + (f[tuple(hi)] - f[tuple(lo)])[cut] / dx[a]
    return np.abs(d)


def divb_rel(bfs, dx):
    """The div B bar in dimensionless form: max over cells of |div B| times
    the smallest cell width, over the largest face |B|.  This is the net
    magnetic flux out of a cell relative to the largest face flux, so a
    value near 1e-16 is one rounding; constrained transport keeps it at
    round-off whatever the field strength and resolution, where an
    absolute |div B| scales with |B| / dx."""
    bmax = max(float(np.abs(f).max()) for f in bfs)
    d = _divb_cells(bfs, dx)
    return float(np.max(d)) * min(dx) / bmax if bmax > 0 else float(np.max(d))


def divb_rel_blocks(spec, blocks):
    """divb_rel over the per-block host arrays of `spec` (any block mesh):
    each block's cells with its own widths, over the largest face |B| of
    the mesh."""
    n = spec.n_cell_block
    nf = bf_faces(n)
    dmax, bmax = 0.0, 0.0
    for b, (_, ba) in enumerate(blocks):
        _, dx = spec.geometry(b)
        f = np.asarray(ba).reshape(bf_host_dims(n)[::-1] + (3,))
        faces = []
        for c in range(3):
            sl = [slice(2, n[a] + 2) for a in range(3)]
            sl[c] = slice(2, nf[c] + 2)
            faces.append(f[sl[2], sl[1], sl[0], c])
        bmax = max(bmax, max(float(np.abs(x).max()) for x in faces))
        act = [dx[a] for a in range(3) if n[a] > 1]
        dmax = max(dmax, float(np.max(_divb_cells(faces, dx))) * min(act))
    return dmax / bmax if bmax > 0 else dmax


@dataclass
class MeshSpec:
    n_cell_global: tuple
    n_cell_block: tuple
    x_min: tuple = (0.0, 0.0, 0.0)
    x_max: tuple = (1.0, 1.0, 1.0)
    bnd: tuple = ("per", "per", "per", "per", "per", "per")

    def __post_init__(self):
        self.n_cell_global = tuple(int(v) for v in self.n_cell_global)
        self.n_cell_block = tuple(int(v) for v in self.n_cell_block)
        for a in range(3):
            if self.n_cell_global[a] % self.n_cell_block[a]:
                raise ValueError("Mesh size indivisible by sub-mesh.")
        self.nblk = tuple(self.n_cell_global[a] // self.n_cell_block[a] for a in range(3))
        self.periodic = tuple(self.bnd[2 * a] == "per" for a in range(3))

    @property
    def n_blocks(self):
        return self.nblk[0] * self.nblk[1] * self.nblk[2]

    def bid(self, r):
        return (r[2] * self.nblk[1] + r[1]) * self.nblk[0] + r[0]

    def reg(self, b):
        nx, ny, _ = self.nblk
        return (b % nx, (b // nx) % ny, b // (nx * ny))

    def geometry(self, b):
        """(xf0, dx0) of block b, as meshgen::set_coord_axes computes them."""
        r = self.reg(b)
        xf0, dx0 = [0.0] * 3, [0.0] * 3
        for a in range(3):
            n = self.n_cell_block[a]

            def loc(i, a=a, n=n):
                if n == 1:
                    return i
                i = (r[a] + i / n) / float(self.nblk[a])
                return self.x_min[a] + i * (self.x_max[a] - self.x_min[a])

            xf0[a] = loc(0.0)
            dx0[a] = loc(1.0) - xf0[a]
        return xf0, dx0

    def neighbours(self, b):
        r = self.reg(b)
        out = [-1] * 27
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    d = (dx, dy, dz)
                    k = ((dz + 1) * 3 + (dy + 1)) * 3 + (dx + 1)
                    q = []
                    ok = True
                    for a in range(3):
                        v = r[a] + d[a]
                        if self.n_cell_block[a] == 1 and d[a] != 0:
                            ok = False
                            break
                        if v < 0 or v >= self.nblk[a]:
                            if not self.periodic[a]:
                                ok = False
                                break
                            v %= self.nblk[a]
                        q.append(v)
                    out[k] = self.bid(q) if ok else -1
        return out

    def block_descs(self, members=1):
        """Descriptors of `members` independent copies of this mesh: copy m
        has block ids m * n_blocks + b and its own time group m."""
        nb = self.n_blocks
        arr = (_capi.KrBlockDesc * (nb * members))()
        for m in range(members):
            for b in range(nb):
                self._desc(arr[m * nb + b], b, m)
        return arr

    def _desc(self, d, b, m):
        nb = self.n_blocks
        d.id = m * nb + b
        d.level = 0
        d.group = m
        r = self.reg(b)
        xf0, dx0 = self.geometry(b)
        for a in range(3):
            d.reg[a] = r[a]
            d.n_ceff[a] = self.n_cell_block[a]
            d.xf0[a] = xf0[a]
            d.dx0[a] = dx0[a]
        for k, v in enumerate(self.neighbours(b)):
            d.nbr[k] = v + m * nb if v >= 0 else -1

    # ---------------------------------------------------- block <-> global
    def block_slices(self, b):
        r = self.reg(b)
        n = self.n_cell_block
        return tuple(slice(r[a] * n[a], (r[a] + 1) * n[a]) for a in range(3))

    def split(self, u, bfs=None):
        """Global fields -> per-block (u_aos, bf_aos) in the reference layout."""
        mhd = bfs is not None
        nvar = 8 if mhd else 5
        n = self.n_cell_block
        out = []
        for b in range(self.n_blocks):
            sx, sy, sz = self.block_slices(b)
            ub = u[:, sz, sy, sx]                       # (5, nz, ny, nx)
            if mhd:
                bx, by, bz = bfs
                bcc = np.stack([
                    0.5 * (bx[sz, sy, sx.start:sx.stop + 1][:, :, :-1] + bx[sz, sy, sx.start:sx.stop + 1][:, :, 1:]),
                    0.5 * (by[sz, sy.start:sy.stop + 1, sx][:, :-1, :] + by[sz, sy.start:sy.stop + 1, sx][:, 1:, :]),
                    0.5 * (bz[sz.start:sz.stop + 1, sy, sx][:-1] + bz[sz.start:sz.stop + 1, sy, sx][1:]),
                ])
                ub = np.concatenate([ub, bcc])
            u_aos = np.ascontiguousarray(np.moveaxis(ub, 0, -1)).reshape(-1)
            bf_aos = None
            if mhd:
                nf = bf_faces(n)
                bf_aos = np.zeros(bf_host_dims(n)[::-1] + (3,))
                r = self.reg(b)
                for c, g in enumerate(bfs):
                    # nf[a] face indices per axis: n + 1 on an active axis (the
                    # transverse +1 slot wraps periodically), 1 on a thin axis
                    idx = []
                    for a in range(3):
                        ii = np.arange(nf[a]) + r[a] * n[a]
                        if a != c:
                            ii = ii % self.n_cell_global[a]
                        idx.append(ii)
                    bf_aos[2:nf[2] + 2, 2:nf[1] + 2, 2:nf[0] + 2, c] = g[np.ix_(idx[2], idx[1], idx[0])]
                bf_aos = bf_aos.reshape(-1)
            out.append((u_aos, bf_aos))
        return out

    def gather(self, blocks, mhd):
        """Per-block (u_aos, bf_aos) -> global (u, (bx, by, bz))."""
        nvar = 8 if mhd else 5
        N = self.n_cell_global
        n = self.n_cell_block
        u = np.zeros((5, N[2], N[1], N[0]))
        bfs = None
        if mhd:
            bfs = (np.zeros((N[2], N[1], N[0] + 1)), np.zeros((N[2], N[1] + 1, N[0])),
                   np.zeros((N[2] + 1, N[1], N[0])))
        for b, (u_aos, bf_aos) in enumerate(blocks):
            sx, sy, sz = self.block_slices(b)
            ub = np.moveaxis(np.asarray(u_aos).reshape(n[2], n[1], n[0], nvar), -1, 0)
            u[:, sz, sy, sx] = ub[:5]
            if mhd:
                f = np.asarray(bf_aos).reshape(bf_host_dims(n)[::-1] + (3,))
                for c in range(3):
                    # component c: n + 1 faces along its axis (on a thin axis
                    # the one face is both sides of the cell), n across
                    sl = [slice(2, n[a] + 2) for a in range(3)]
                    dst = list(self.block_slices(b))
                    if n[c] > 1:
                        sl[c] = slice(2, n[c] + 3)
                        dst[c] = slice(dst[c].start, dst[c].stop + 1)
                    v = f[sl[2], sl[1], sl[0], c]
                    bfs[c][dst[2], dst[1], dst[0]] = v
                    if n[c] == 1:
                        dst[c] = slice(dst[c].start + 1, dst[c].stop + 1)
                        bfs[c][dst[2], dst[1], dst[0]] = v
        return u, bfs


class RefinedMeshSpec:
    """A statically refined block mesh (SMR): blocks of one size at several
    levels, the reference's tree_t built from ``[refine_region_*]`` entries.

    ``refine`` is a sequence of ``(level, x_min, x_max)`` boxes; every block
    of that level whose index lies in the box (inclusive,
    meshgen.cpp:111-158 with ``search_blk_idx`` at :355-374) is inserted
    with tree_t::insert (tree.cpp:185-205): its parent is masked, the
    parent's children and the parent's 26 same-level neighbours are
    inserted too, which keeps the mesh 2:1 balanced across faces, edges and
    corners.  ``fill_base_level`` (tree.cpp:207-218) completes level 0.
    Blocks are the unmasked tree nodes in the tree's order: level, then
    x, y, z of the logical index (region.h lex_ord_t).
    """

    def __init__(self, n_cell_global, n_cell_block, refine=(), x_min=(0.0, 0.0, 0.0),
                 x_max=(1.0, 1.0, 1.0), bnd=("per",) * 6):
        self.n_cell_global = tuple(int(v) for v in n_cell_global)
        self.n_cell_block = tuple(int(v) for v in n_cell_block)
        self.x_min, self.x_max, self.bnd = tuple(x_min), tuple(x_max), tuple(bnd)
        for a in range(3):
            if self.n_cell_global[a] % self.n_cell_block[a]:
                raise ValueError("Mesh size indivisible by sub-mesh.")
        self.nblk = tuple(self.n_cell_global[a] // self.n_cell_block[a] for a in range(3))
        self.periodic = tuple(self.bnd[2 * a] == "per" for a in range(3))
        self.refine = [(int(l), tuple(lo), tuple(hi)) for l, lo, hi in refine]
        tree = {}

        def lim(level, a):
            return self.nblk[a] << level

        def regularize(level, r):
            return tuple((r[a] + lim(level, a)) % lim(level, a) if self.periodic[a] else r[a]
                         for a in range(3))

        def inside(level, r):
            return all(0 <= r[a] < lim(level, a) for a in range(3))

        def insert(level, r):
            if (level,) + r in tree:
                return
            if level > 0:
                rc = tuple(v // 2 for v in r)
                insert(level - 1, rc)
                tree[(level - 1,) + rc] = True
                for c in range(8):
                    tree.setdefault((level,) + tuple(2 * rc[a] + ((c >> a) & 1) for a in range(3)),
                                    False)
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            if dx == dy == dz == 0:
                                continue
                            q = regularize(level - 1, (rc[0] + dx, rc[1] + dy, rc[2] + dz))
                            if inside(level - 1, q):
                                insert(level - 1, q)
            tree.setdefault((level,) + r, False)

        for level, lo, hi in self.refine:
            if level <= 0:
                continue
            i0 = [self.search_blk_idx(lo[a], level, a) for a in range(3)]
            i1 = [self.search_blk_idx(hi[a], level, a) for a in range(3)]
            if min(i0 + i1) < 0:
                raise ValueError("Incorrect SMR region")
            for z in range(i0[2], i1[2] + 1):
                for y in range(i0[1], i1[1] + 1):
                    for x in range(i0[0], i1[0] + 1):
                        insert(level, (x, y, z))
        for z in range(self.nblk[2]):
            for y in range(self.nblk[1]):
                for x in range(self.nblk[0]):
                    tree.setdefault((0, x, y, z), False)
        self.tree = dict(tree)               # (level, x, y, z) -> masked
        self.leaves = [k for k in sorted(tree) if not tree[k]]
        self.n_levels = 1 + max(k[0] for k in self.leaves)

    def location(self, x_logic, a):
        return self.x_min[a] + x_logic * (self.x_max[a] - self.x_min[a])

    def search_blk_idx(self, loc, lvl, a):
        """Block index of level `lvl` along axis `a` containing `loc`."""
        if loc > self.x_max[a] or loc < self.x_min[a]:
            return -1
        i_tot = self.nblk[a] << lvl
        i_l, i_r = 0, i_tot
        while i_r - i_l > 1:
            i_m = (i_l + i_r) // 2
            if self.location(i_m / float(i_tot), a) < loc:
                i_l = i_m
            else:
                i_r = i_m
        return i_l

    @property
    def n_blocks(self):
        return len(self.leaves)

    @property
    def n_cells(self):
        n = self.n_cell_block
        return self.n_blocks * n[0] * n[1] * n[2]

    def level(self, b):
        return self.leaves[b][0]

    def reg(self, b):
        return self.leaves[b][1:]

    def geometry(self, b):
        """(xf0, dx0) of block b (meshgen.cpp:286-321 at the block's level)."""
        level, r = self.level(b), self.reg(b)
        xf0, dx0 = [0.0] * 3, [0.0] * 3
        for a in range(3):
            n = self.n_cell_block[a]

            def loc(i, a=a, n=n):
                i = (r[a] + i / n) / float(self.nblk[a] << level)
                return self.location(i, a)

            xf0[a] = loc(0.0)
            dx0[a] = loc(1.0) - xf0[a]
        return xf0, dx0

    def block_descs(self, members=1):
        if members != 1:
            raise ValueError("refined meshes have one member")
        arr = (_capi.KrBlockDesc * self.n_blocks)()
        for b in range(self.n_blocks):
            d = arr[b]
            d.id, d.level, d.group = b, self.level(b), 0
            xf0, dx0 = self.geometry(b)
            for a in range(3):
                d.reg[a] = self.reg(b)[a]
                d.n_ceff[a] = self.n_cell_block[a]
                d.xf0[a], d.dx0[a] = xf0[a], dx0[a]
            nb = self.neighbours(b)
            for k in range(27):
                d.nbr[k] = nb[k]
        return arr

    def neighbours(self, b):
        """Same-level neighbour per direction (-1 where the neighbour region
        is at another level or outside a non-periodic side); the library
        derives the cross-level topology itself from level and reg."""
        if not hasattr(self, "_where"):
            self._where = {leaf: i for i, leaf in enumerate(self.leaves)}
        L, r = self.level(b), self.reg(b)
        out = [-1] * 27
        for k in range(27):
            d = (k % 3 - 1, (k // 3) % 3 - 1, k // 9 - 1)
            q = []
            for a in range(3):
                v = r[a] + d[a]
                lim = self.nblk[a] << L
                if v < 0 or v >= lim:
                    if not self.periodic[a]:
                        break
                    v %= lim
                q.append(v)
            else:
                out[k] = self._where.get((L,) + tuple(q), -1)
        return out

    def split(self, u, bfs=None):
        raise TypeError("refined meshes take per-block arrays (upload_blocks)")


class MergedMeshSpec:
    """A refined mesh given with blocks of two sizes: every group of eight
    leaves ``reg + {0,1}^3`` of one level, ``reg`` even on each axis, is one
    block of twice the block size (the eight's level, ``reg`` the lowest
    one's index, the lowest one's corner and cell), the other leaves stay
    blocks of the base size.  Blocks follow the base mesh's leaf order, a
    merged block at its lowest leaf's place; ``parts[b]`` lists the base
    leaves of block b (octant o = x + 2 y + 4 z bits).  The reference has
    one block size per mesh (mesh/meshgen/meshgen.cpp:446-448); the library
    runs a merged block as its eight leaves, so the cells are the base
    mesh's."""

    def __init__(self, base):
        self.base = base
        self.n_cell_block = base.n_cell_block
        self.x_min, self.x_max, self.bnd = base.x_min, base.x_max, base.bnd
        self.periodic = base.periodic
        self.n_levels = base.n_levels
        where = {leaf: i for i, leaf in enumerate(base.leaves)}
        used, self.parts = set(), []
        for i, leaf in enumerate(base.leaves):
            if i in used:
                continue
            L, r = leaf[0], leaf[1:]
            kids = None
            if all(v % 2 == 0 for v in r):
                kids = [where.get((L,) + tuple(r[a] + ((o >> a) & 1) for a in range(3)))
                        for o in range(8)]
                if any(k is None for k in kids):
                    kids = None
            if kids:
                used.update(kids)
                self.parts.append(kids)
            else:
                self.parts.append([i])

    @property
    def n_blocks(self):
        return len(self.parts)

    @property
    def n_cells(self):
        return self.base.n_cells

    def merged(self, b):
        return len(self.parts[b]) == 8

    def block_cells(self, b):
        f = 2 if self.merged(b) else 1
        return tuple(f * v for v in self.n_cell_block)

    def block_descs(self, members=1):
        if members != 1:
            raise ValueError("refined meshes have one member")
        arr = (_capi.KrBlockDesc * self.n_blocks)()
        for b in range(self.n_blocks):
            d, lo = arr[b], self.parts[b][0]
            d.id, d.level, d.group = b, self.base.level(lo), 0
            xf0, dx0 = self.base.geometry(lo)
            nb = self.block_cells(b)
            for a in range(3):
                d.reg[a] = self.base.reg(lo)[a]
                d.n_ceff[a] = nb[a]
                d.xf0[a], d.dx0[a] = xf0[a], dx0[a]
            for k in range(27):
                d.nbr[k] = -1          # derived from level and reg
        return arr

    def _octant(self, o):
        n = self.n_cell_block
        return [((o >> a) & 1) * n[a] for a in range(3)]

    def merge_blocks(self, blocks):
        """Per-leaf (u_aos, bf_aos) of the base mesh -> per-block arrays;
        a face two leaves share is taken from the upper leaf."""
        n = self.n_cell_block
        out = []
        for parts in self.parts:
            if len(parts) == 1:
                out.append(blocks[parts[0]])
                continue
            nv = np.asarray(blocks[parts[0]][0]).size // (n[0] * n[1] * n[2])
            U = np.zeros((2 * n[2], 2 * n[1], 2 * n[0], nv))
            mhd = blocks[parts[0]][1] is not None
            B = np.zeros((2 * n[2] + 5, 2 * n[1] + 5, 2 * n[0] + 5, 3)) if mhd else None
            for o, leaf in enumerate(parts):
                h = self._octant(o)
                U[h[2]:h[2] + n[2], h[1]:h[1] + n[1], h[0]:h[0] + n[0]] = \
                    np.asarray(blocks[leaf][0]).reshape(n[2], n[1], n[0], nv)
                if mhd:
                    bo = np.asarray(blocks[leaf][1]).reshape(n[2] + 5, n[1] + 5, n[0] + 5, 3)
                    B[2 + h[2]:3 + h[2] + n[2], 2 + h[1]:3 + h[1] + n[1], 2 + h[0]:3 + h[0] + n[0]] = \
                        bo[2:3 + n[2], 2:3 + n[1], 2:3 + n[0]]
            out.append((U.ravel(), B.ravel() if mhd else None))
        return out

    def split_blocks(self, blocks):
        """Per-block arrays -> per-leaf arrays of the base mesh (zero ghost
        layers, as kr_download gives them)."""
        n = self.n_cell_block
        out = [None] * self.base.n_blocks
        for b, parts in enumerate(self.parts):
            if len(parts) == 1:
                out[parts[0]] = blocks[b]
                continue
            ua, ba = blocks[b]
            nv = np.asarray(ua).size // (8 * n[0] * n[1] * n[2])
            U = np.asarray(ua).reshape(2 * n[2], 2 * n[1], 2 * n[0], nv)
            B = None if ba is None else np.asarray(ba).reshape(2 * n[2] + 5, 2 * n[1] + 5, 2 * n[0] + 5, 3)
            for o, leaf in enumerate(parts):
                h = self._octant(o)
                u = U[h[2]:h[2] + n[2], h[1]:h[1] + n[1], h[0]:h[0] + n[0]].copy()
                bo = None
                if B is not None:
                    bo = np.zeros((n[2] + 5, n[1] + 5, n[0] + 5, 3))
                    bo[2:3 + n[2], 2:3 + n[1], 2:3 + n[0]] = \
                        B[2 + h[2]:3 + h[2] + n[2], 2 + h[1]:3 + h[1] + n[1], 2 + h[0]:3 + h[0] + n[0]]
                    bo = bo.ravel()
                out[leaf] = (u.ravel(), bo)
        return out


@dataclass
class Physics:
    mhd: bool = True
    gamma: float = 5.0 / 3.0
    cfl: float = 0.3
    rho_floor: float = 1e-9
    tmp_floor: float = 1e-9
    t_0: float = 0.0
    t_lim: float = 3.4028234663852886e38
    dt_init: float = 0.0
    dt_expand: float = 2.0
    top_up: bool = False
    # ghost state of "fixd" (Dirichlet) sides: primitives rho, p, v0, v1, v2
    # and B -- the reference's [init] rho0 / pre0 / vel0 / b0
    fix_w: tuple = (1.0, 1.0, 0.0, 0.0, 0.0)
    fix_b: tuple = (0.0, 0.0, 0.0)


class Context:
    """One device context of the C-ABI over a uniform block mesh."""

    def __init__(self, spec: MeshSpec, phys: Physics, precision="fp64", path="auto",
                 tile=(0, 0, 0), members=1):
        """`members` > 1 makes an ensemble: that many independent copies of
        the mesh in one context, each its own time group (own t and dt),
        all advanced by the same launches.  upload / download / time take
        the member index."""
        lib = _capi.load()
        self.lib, self.spec, self.phys = lib, spec, phys
        self.members = int(members)
        self.mhd = phys.mhd
        cfg = _capi.KrConfig()
        cfg.model = _capi.KR_MHD if phys.mhd else _capi.KR_HYDRO
        cfg.precision = {"fp64": _capi.KR_FP64, "mixed": _capi.KR_MIXED}[precision]
        cfg.gamma, cfg.cfl = phys.gamma, phys.cfl
        cfg.rho_floor, cfg.tmp_floor = phys.rho_floor, phys.tmp_floor
        for s in range(6):
            cfg.bnd[s] = _capi.BND_CODES[spec.bnd[s]]
        cfg.t_0, cfg.t_lim = phys.t_0, phys.t_lim
        cfg.dt_init, cfg.dt_expand = phys.dt_init, phys.dt_expand
        cfg.top_up = int(phys.top_up)
        for a in range(3):
            cfg.n_cell_block[a] = spec.n_cell_block[a]
        cfg.kernel_path = _capi.PATHS[path]
        tile = tuple(tile) + (0,) * (3 - len(tile))
        for k in range(3):
            cfg.tile[k] = int(tile[k])
        for k in range(5):
            cfg.fix_w[k] = float(phys.fix_w[k])
        for k in range(3):
            cfg.fix_b[k] = float(phys.fix_b[k])
        h = C.c_void_p()
        rc = lib.kr_create(C.byref(cfg), C.byref(h))
        _capi.check(rc, None, "kr_create")
        self.h = h
        descs = spec.block_descs(self.members)
        _capi.check(lib.kr_add_blocks(h, spec.n_blocks * self.members, descs), h,
                    "kr_add_blocks")

    def close(self):
        if self.h:
            self.lib.kr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, u, bfs=None, member=0):
        off = member * self.spec.n_blocks
        for b, (ua, ba) in enumerate(self.spec.split(u, bfs if self.mhd else None)):
            ua = np.ascontiguousarray(ua, dtype=np.float64)
            bp = None
            if ba is not None:
                ba = np.ascontiguousarray(ba, dtype=np.float64)
                bp = ba.ctypes.data
            _capi.check(self.lib.kr_upload(self.h, off + b, ua.ctypes.data, bp), self.h,
                        "kr_upload")

    def download(self, member=0):
        off = member * self.spec.n_blocks
        n = self.spec.n_cell_block
        nvar = 8 if self.mhd else 5
        blocks = []
        for b in range(self.spec.n_blocks):
            ua = np.zeros(n[0] * n[1] * n[2] * nvar)
            ba = np.zeros(bf_host_size(n)) if self.mhd else None
            _capi.check(self.lib.kr_download(self.h, off + b, ua.ctypes.data,
                                             ba.ctypes.data if ba is not None else None),
                        self.h, "kr_download")
            blocks.append((ua, ba))
        return self.spec.gather(blocks, self.mhd)

    def upload_blocks(self, blocks):
        """Per-block (u_aos, bf_aos) in the reference's host layout, in
        block order (the only way to fill a refined mesh)."""
        for b, (ua, ba) in enumerate(blocks):
            ua = np.ascontiguousarray(ua, dtype=np.float64)
            bp = None
            if self.mhd:
                ba = np.ascontiguousarray(ba, dtype=np.float64)
                bp = ba.ctypes.data
            _capi.check(self.lib.kr_upload(self.h, b, ua.ctypes.data, bp), self.h, "kr_upload")

    def download_blocks(self):
        nvar = 8 if self.mhd else 5
        out = []
        for b in range(self.spec.n_blocks):
            n = self.spec.block_cells(b) if hasattr(self.spec, "block_cells") else self.spec.n_cell_block
            ua = np.zeros(n[0] * n[1] * n[2] * nvar)
            ba = np.zeros(bf_host_size(n)) if self.mhd else None
            _capi.check(self.lib.kr_download(self.h, b, ua.ctypes.data,
                                             ba.ctypes.data if ba is not None else None),
                        self.h, "kr_download")
            out.append((ua, ba))
        return out

    def step(self):
        dt = C.c_double()
        _capi.check(self.lib.kr_step(self.h, C.byref(dt)), self.h, "kr_step")
        return dt.value

    def advance(self, n_cycles):
        done = C.c_int()
        _capi.check(self.lib.kr_advance(self.h, int(n_cycles), C.byref(done)), self.h,
                    "kr_advance")
        return done.value

    def time(self, member=None):
        """(t, dt_last, dt_seed) of the mesh, or (t, dt_last, cycles) of
        ensemble member `member`."""
        if member is not None:
            t, dt, cyc = C.c_double(), C.c_double(), C.c_int()
            _capi.check(self.lib.kr_get_group_time(self.h, int(member), C.byref(t), C.byref(dt),
                                                   C.byref(cyc)), self.h, "kr_get_group_time")
            return t.value, dt.value, cyc.value
        t, dt, seed = C.c_double(), C.c_double(), C.c_double()
        _capi.check(self.lib.kr_get_time(self.h, C.byref(t), C.byref(dt), C.byref(seed)),
                    self.h, "kr_get_time")
        return t.value, dt.value, seed.value

    def set_time(self, t, dt_seed):
        """Time of every group and the dt seed of the next cycle (kr_set_time)."""
        _capi.check(self.lib.kr_set_time(self.h, float(t), float(dt_seed)), self.h, "kr_set_time")

    def divb_max(self):
        v = C.c_double()
        _capi.check(self.lib.kr_divb_max(self.h, C.byref(v)), self.h, "kr_divb_max")
        return v.value

    def profile(self, n_cycles):
        _capi.check(self.lib.kr_profile(self.h, int(n_cycles)), self.h, "kr_profile")
        return self.stats()

    def stats(self):
        s = _capi.KrStats()
        _capi.check(self.lib.kr_get_stats(self.h, C.byref(s)), self.h, "kr_get_stats")
        return {
            "gpu_ms": s.gpu_ms, "cycles": s.cycles, "n_blocks": s.n_blocks,
            "n_cells": s.n_cells,
            "kernel_ms": {k: s.kernel_ms[i] for i, k in enumerate(_capi.KERNEL_KINDS)},
            "kernel_calls": {k: s.kernel_calls[i] for i, k in enumerate(_capi.KERNEL_KINDS)},
            "alg_bytes_per_cell_update": s.alg_bytes_per_cell_update,
            "sweep_tile": tuple(s.sweep_tile),
            "sweep_kernel": s.sweep_kernel,
            "tune_ms": tuple(s.tune_ms),
            "tune_pending": s.tune_pending,
            "graph_ready": bool(s.graph_ready),
        }
