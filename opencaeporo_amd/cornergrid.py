"""Corner-point geometry for input decks (COORD / ZCORN): cell volumes,
depths and sizes, and the connections between cells, including the
non-neighbour connections across faults, with their half-transmissibility
factors Ad_dd, as the reference's OCP_COORD::SetupCornerPoints computes them
(src/CornerGrid.cpp:770-1747), then BulkConn::CalAkd's harmonic
transmissibilities (src/BulkConn.cpp:89-117).  Host-side, numpy; run once
when a deck is read.

Cells are indexed n = k nx ny + j nx + i.  A cell's corners 0-3 are its top
face at pillars (i, j), (i+1, j), (i+1, j+1), (i, j+1); 4-7 its bottom face
at the same pillars.  A corner lies on its pillar (the straight line through
the pillar's top and bottom COORD points) at the ZCORN depth.

Connections.  Across each lateral face (x-, x+, y-, y+) a cell meets the
neighbouring column's cell of the same layer and, where the faces are offset
(a fault), the cells above and below it for as long as their faces still
reach this face.  The overlap of two faces is the face with its corners
pulled onto the other face's edges when the overlap is a quadrilateral on
the shared pillars, and otherwise the overlap area of the two faces
projected onto the plane of their pillars, along this face's normal (the
reference's twelve-case construction, kept case for case because the
transmissibilities of a drop-in must be the reference's).  Vertical faces
connect layers k and k +- 1 of a column.  A half connection of cell n to m
carries Ad_dd = (A . d) / (d . d), A the overlap's area vector and d the
vector from n's centre to the centre of n's face; a connection needs both
halves, both positive.
"""
import numpy as np

TEENY = 1e-3        # depth tolerance of the face-offset tests
SMALL_REAL = 1e-10  # determinant below which two edges count as parallel


class CornerError(ValueError):
    pass


def corner_points(nx, ny, nz, coord, zcorn):
    """(nb, 8, 3) corner points from COORD ((nx+1)(ny+1) pillars x 6, deck
    order) and ZCORN (8 nx ny nz, deck order)."""
    coord = np.asarray(coord, dtype=np.float64)
    zcorn = np.asarray(zcorn, dtype=np.float64)
    if coord.size != 6 * (nx + 1) * (ny + 1):
        raise CornerError(f"COORD needs {6 * (nx + 1) * (ny + 1)} values, has {coord.size}")
    if zcorn.size != 8 * nx * ny * nz:
        raise CornerError(f"ZCORN needs {8 * nx * ny * nz} values, has {zcorn.size}")
    pil = coord.reshape(ny + 1, nx + 1, 2, 3)  # [J][I][top/bottom][xyz]
    # ZCORN: per layer a top then a bottom sheet of (2 ny) x (2 nx) depths
    z = zcorn.reshape(nz, 2, ny, 2, nx, 2)  # [k][top/bot][j][near/far in y][i][near/far in x]
    # corner c of the hexahedron: (sheet, y side, x side)
    sides = [(0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), (1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)]
    nb = nx * ny * nz
    P = np.empty((nz, ny, nx, 8, 3))
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    for c, (s, ys, xs) in enumerate(sides):
        zc = z[:, s, :, ys, :, xs]  # (nz, ny, nx)
        p = pil[jj + ys, ii + xs]    # (ny, nx, 2, 3)
        top, bot = p[:, :, 0, :], p[:, :, 1, :]
        dzp = bot[:, :, 2] - top[:, :, 2]
        if np.any(dzp == 0):
            raise CornerError("a pillar has equal top and bottom depths")
        # the point of the pillar at depth zc
        f = (bot[None, :, :, 2] - zc) / dzp[None]
        P[..., c, 0] = bot[None, :, :, 0] - f * (bot[None, :, :, 0] - top[None, :, :, 0])
        P[..., c, 1] = bot[None, :, :, 1] - f * (bot[None, :, :, 1] - top[None, :, :, 1])
        P[..., c, 2] = zc
    return P.reshape(nb, 8, 3)


# 2-point Gauss-Legendre abscissae on [0, 1]: exact for the Jacobian
# determinant of a trilinear hexahedron (degree <= 2 in each coordinate)
_G = np.array([0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)])


def hex_volume(P):
    """Volume of each trilinear hexahedron (corners as above): the integral
    of the map's Jacobian determinant over the unit cube."""
    # trilinear map x(u, v, w) with corner c at (u, v, w) = (xs, ys, sheet)
    uvw = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)],
                   dtype=np.float64)
    vol = np.zeros(P.shape[0])
    for u in _G:
        for v in _G:
            for w in _G:
                q = np.array([u, v, w])
                # d(shape)/d(u, v, w) of the eight trilinear shape functions
                d = np.empty((8, 3))
                for c in range(8):
                    a = np.where(uvw[c] == 1, q, 1 - q)
                    s = np.where(uvw[c] == 1, 1.0, -1.0)
                    d[c] = [s[0] * a[1] * a[2], a[0] * s[1] * a[2], a[0] * a[1] * s[2]]
                J = np.einsum("ncx,cy->nxy", P, d)  # (nb, 3 coords, 3 params)
                vol += np.linalg.det(J) / 8.0
    return np.abs(vol)


def face_vector(f):
    """Area vector of quadrilaterals f (..., 4, 3): the mean of the two
    triangles' cross products about the diagonal p1-p3 (VectorFace)."""
    p0, p1, p2, p3 = f[..., 0, :], f[..., 1, :], f[..., 2, :], f[..., 3, :]
    return 0.5 * (np.cross(p2 - p1, p0 - p1) + np.cross(p0 - p3, p2 - p3))


def _cross_point(a0, a1, b0, b1):
    """Intersection of the lines a0-a1 and b0-b1 (2D); a0 when parallel."""
    a11, a12 = a1[1] - a0[1], a0[0] - a1[0]
    a21, a22 = b1[1] - b0[1], b0[0] - b1[0]
    r1 = a11 * a0[0] + a12 * a0[1]
    r2 = a21 * b0[0] + a22 * b0[1]
    det = a11 * a22 - a12 * a21
    if abs(det) > SMALL_REAL:
        return np.array([(r1 * a22 - r2 * a12) / det, (a11 * r2 - a21 * r1) / det])
    return np.array(a0, dtype=np.float64)


def _quad_area(q):
    """|z| of the area vector of a planar quadrilateral given by 4 2D points."""
    p = [np.array([x[0], x[1], 0.0]) for x in q]
    return abs(face_vector(np.array(p))[2])


def overlap_area(F1, F2):
    """Common area of two quadrilaterals on the same pair of pillars,
    projected to 2D (second coordinate the depth).  Points 0, 1 are the
    upper edge, 2, 3 the lower edge (CalAreaNotQuadr,
    src/CornerGrid.cpp:198-565, case for case)."""
    up1, lo1, up2, lo2 = (F1[0], F1[1]), (F1[2], F1[3]), (F2[0], F2[1]), (F2[2], F2[3])
    cps, code = [], 0
    for bit, (e1, e2) in enumerate(((up1, up2), (lo1, up2), (up1, lo2), (lo1, lo2))):
        cp = _cross_point(e1[0], e1[1], e2[0], e2[1])
        cps.append(cp)
        if (cp[0] - e1[0][0]) * (cp[0] - e1[1][0]) < 0:
            code |= 1 << bit
    c0, c1, c2, c3 = cps
    A0, A1, A2, A3 = F1
    B0, B1, B2, B3 = F2
    deeper = lambda a, b: a if a[1] > b[1] else b
    shallower = lambda a, b: b if a[1] > b[1] else a
    tri = lambda a, b, c: 0.5 * abs((a[0] - c[0]) * (b[1] - c[1]) - (a[1] - c[1]) * (b[0] - c[0]))
    mid = lambda a, b: 0.5 * (np.asarray(a) + np.asarray(b))
    if code in (1, 8, 9):
        # the upper edges cross (1), the lower ones (8), or both (9): two
        # quadrilaterals either side of the crossing(s)
        t0, t1p = deeper(A0, B0), deeper(A1, B1)
        t3, t2p = shallower(A3, B3), shallower(A2, B2)
        if code == 1:
            m = mid(t3, t2p)
            return _quad_area([t0, c0, m, t3]) + _quad_area([c0, t1p, t2p, m])
        if code == 8:
            m = mid(t0, t1p)
            return _quad_area([t0, m, c3, t3]) + _quad_area([m, t1p, t2p, c3])
        return _quad_area([t0, c0, c3, t3]) + _quad_area([c0, t1p, t2p, c3])
    if code == 2:
        a, b = (A3, B0) if A3[1] > B0[1] else (A2, B1)
        return tri(a, b, c1)
    if code == 3:
        a, b = (A2, A1) if A0[1] < B0[1] else (A3, A0)
        return _quad_area([c0, c1, a, b])
    if code == 4:
        a, b = (A0, B3) if A0[1] < B3[1] else (A1, B2)
        return tri(a, b, c2)
    if code == 5:
        a, b = (B3, B0) if B3[1] > A0[1] else (B2, B1)
        return _quad_area([c0, c2, a, b])
    if code == 10:
        a, b = (B1, B2) if A2[1] > B1[1] else (B3, B0)
        return _quad_area([c1, c3, a, b])
    if code == 11:
        q1p1, q1p2 = (A1, B2) if A0[1] < B0[1] else (A0, B3)
        m = mid(c0, q1p1)
        return _quad_area([c0, m, c3, c1]) + _quad_area([m, q1p1, q1p2, c3])
    if code == 12:
        a, b = (A3, A0) if A2[1] > B2[1] else (A2, A1)
        return _quad_area([c2, c3, a, b])
    if code == 13:
        p0, p3 = (B0, A3) if A2[1] > B2[1] else (B1, A2)
        m = mid(p0, c0)
        return _quad_area([p0, m, c3, p3]) + _quad_area([m, c0, c2, c3])
    if code == 15:
        return _quad_area([c0, c2, c3, c1])
    return 0.0


# Lateral faces: (corners of this cell's face, corners of the neighbour's
# face, neighbour's column offset (di, dj), direction, projection axes of
# the overlap, order of the projected points, axis of the area correction)
_LATERAL = (
    ((0, 4, 7, 3), (1, 5, 6, 2), (-1, 0), 1),   # x-
    ((2, 6, 5, 1), (3, 7, 4, 0), (1, 0), 1),    # x+
    ((1, 5, 4, 0), (2, 6, 7, 3), (0, -1), 2),   # y-
    ((3, 7, 6, 2), (0, 4, 5, 1), (0, 1), 2),    # y+
)


class CornerGrid:
    """Geometry of a corner-point grid: nb, v, depth, dx, dy, dz (cell
    centre to face centre, both sides), centre (nb, 3) and the connection
    list (begin < end in the reference's OCP_COORD order) with direction
    (1 x, 2 y, 3 z) and the two Ad_dd factors."""

    def __init__(self, nx, ny, nz, coord, zcorn):
        self.nx, self.ny, self.nz = nx, ny, nz
        nb = nx * ny * nz
        self.nb = nb
        P = corner_points(nx, ny, nz, coord, zcorn)
        self.corners = P
        self.v = hex_volume(P)
        self.center = P.mean(axis=1)
        self.depth = self.center[:, 2].copy()
        pil = np.asarray(coord, dtype=np.float64).reshape(ny + 1, nx + 1, 2, 3)
        # blocks placed along y+ (1) or y- (-1): the sign of the area vectors
        self.forward = 1.0 if pil[1, 0, 0, 1] > pil[0, 0, 0, 1] else -1.0
        half = []  # (cell, neighbour, direction, Ad_dd, order)
        d_face = {}
        for slot, (fc, oc, (di, dj), direction) in enumerate(_LATERAL):
            d_face[slot] = self._lateral(P, fc, oc, di, dj, direction, slot, half)
        # vertical faces: z- (top, corners 0 3 2 1) and z+ (bottom, 5 6 7 4)
        idx = np.arange(nb)
        k = idx // (nx * ny)
        for slot, (fc, dk) in ((4, ((0, 3, 2, 1), -1)), (5, ((5, 6, 7, 4), 1))):
            F = P[:, fc, :]
            d = F.mean(axis=1) - self.center
            d_face[slot] = d
            ok = (k + dk >= 0) & (k + dk < nz)
            A = face_vector(F[ok])
            dd = d[ok]
            ad = np.einsum("ij,ij->i", A, dd) / np.einsum("ij,ij->i", dd, dd) * self.forward
            half.append((idx[ok], idx[ok] + dk * nx * ny, np.full(ok.sum(), 3), ad,
                         np.full(ok.sum(), slot * 1_000_000)))
        self.dx = np.linalg.norm(d_face[1] - d_face[0], axis=1)
        self.dy = np.linalg.norm(d_face[3] - d_face[2], axis=1)
        self.dz = np.linalg.norm(d_face[5] - d_face[4], axis=1)
        self._pair(half)

    def _lateral(self, P, fc, oc, di, dj, direction, slot, half):
        nx, ny, nz = self.nx, self.ny, self.nz
        nxny = nx * ny
        idx = np.arange(self.nb)
        i, j, k = idx % nx, (idx // nx) % ny, idx // nxny
        F = P[:, fc, :]
        d = F.mean(axis=1) - self.center
        has = (i + di >= 0) & (i + di < nx) & (j + dj >= 0) & (j + dj < ny)
        cells = idx[has]
        base = cells + di + dj * nx  # the neighbour column's cell in layer k

        def visit(c, o, step):
            """Face offsets of cells c against cells o; records the half
            connections that overlap; returns the up / down continuation."""
            Fc, Oc = F[c], P[o][:, oc, :]
            z, oz = Fc[:, :, 2], Oc[:, :, 2]
            f0 = np.where(oz[:, 0] > z[:, 0] + TEENY, 1, np.where(oz[:, 0] < z[:, 0] - TEENY, -1, 0))
            f1 = np.where(oz[:, 1] > z[:, 1] + TEENY, 1, np.where(oz[:, 1] < z[:, 1] - TEENY, -1, 0))
            f2 = np.where(oz[:, 2] > z[:, 2] + TEENY, 1, np.where(oz[:, 2] < z[:, 2] - TEENY, -1, 0))
            f3 = np.where(oz[:, 3] > z[:, 3] + TEENY, 1, np.where(oz[:, 3] < z[:, 3] - TEENY, -1, 0))
            jump = (((oz[:, 1] <= z[:, 0]) & (oz[:, 2] <= z[:, 3])) |
                    ((oz[:, 0] >= z[:, 1]) & (oz[:, 3] >= z[:, 2])))
            quad = ((f0 * f3 >= 0) & (oz[:, 0] <= z[:, 1]) & (oz[:, 3] <= z[:, 2]) & (f1 * f2 >= 0) &
                    (oz[:, 1] >= z[:, 0]) & (oz[:, 2] >= z[:, 3]))
            # quadrilateral overlap: this face with corners pulled onto the other's
            T = Fc.copy()
            T[:, 0] = np.where((f0 > 0)[:, None], Oc[:, 0], Fc[:, 0])
            T[:, 1] = np.where((f1 < 0)[:, None], Oc[:, 1], Fc[:, 1])
            T[:, 2] = np.where((f2 < 0)[:, None], Oc[:, 2], Fc[:, 2])
            T[:, 3] = np.where((f3 > 0)[:, None], Oc[:, 3], Fc[:, 3])
            A = face_vector(T)
            ax = direction - 1       # the face's normal axis (x or y)
            other = 1 - ax           # the horizontal axis along the face
            for r in np.flatnonzero(~jump & ~quad):
                # the projected overlap, this face's normal rescaled to it
                if direction == 1:
                    order = (3, 0, 1, 2)
                else:
                    order = (0, 3, 2, 1)
                fp = [(Fc[r, q, other], Fc[r, q, 2]) for q in order]
                op = [(Oc[r, q, other], Oc[r, q, 2]) for q in order]
                area = overlap_area(np.array(fp), np.array(op))
                a = face_vector(Fc[r]).copy()
                if abs(a[ax]) >= 1e-6:
                    n_ax = abs(a[ax])
                    a[other] = a[other] / n_ax * area
                    a[2] = a[2] / n_ax * area
                    a[ax] = (1.0 if a[ax] >= 0 else -1.0) * area
                A[r] = a
            keep = ~jump
            dd = d[c[keep]]
            ad = np.einsum("ij,ij->i", A[keep], dd) / np.einsum("ij,ij->i", dd, dd) * self.forward
            half.append((c[keep], o[keep], np.full(keep.sum(), direction), ad,
                         np.full(keep.sum(), slot * 1_000_000 + step)))
            return (f0 > 0) | (f3 > 0), (f1 < 0) | (f2 < 0)

        up, down = visit(cells, base, 0)
        # the other column's cells above, then below, while they still reach
        for sgn, go, order0 in ((-1, up, 1), (1, down, 500_000)):
            c, o, kk, cont = cells, base, k[has], go
            step = 1
            while True:
                ok = cont & (kk + sgn * step >= 0) & (kk + sgn * step < nz)
                if not ok.any():
                    break
                c, o, kk = c[ok], o[ok], kk[ok]
                u, dn = visit(c, o + sgn * step * nxny, order0 + step)
                cont = u if sgn < 0 else dn
                step += 1
        return d

    def _pair(self, half):
        """OCP_COORD's pairing: each half connection n -> m (m > n) with the
        first half connection m -> n, both Ad_dd > 0; in the order of n,
        then of n's half connections."""
        cell = np.concatenate([h[0] for h in half]).astype(np.int64)
        nbr = np.concatenate([h[1] for h in half]).astype(np.int64)
        dirn = np.concatenate([h[2] for h in half]).astype(np.int32)
        ad = np.concatenate([h[3] for h in half])
        seq = np.concatenate([h[4] for h in half]).astype(np.int64)
        # each cell's half connections in the order the reference adds them
        o = np.lexsort((seq, cell))
        cell, nbr, dirn, ad, seq = cell[o], nbr[o], dirn[o], ad[o], seq[o]
        nb = np.int64(self.nb)
        key = cell * nb + nbr
        # the first half connection of each (cell, neighbour): lowest seq
        by_key = np.lexsort((seq, key))
        ukeys, first_pos = np.unique(key[by_key], return_index=True)
        first = by_key[first_pos]
        fwd = np.flatnonzero(nbr > cell)
        want = nbr[fwd] * nb + cell[fwd]
        pos = np.searchsorted(ukeys, want)
        found = (pos < len(ukeys)) & (ukeys[np.minimum(pos, len(ukeys) - 1)] == want)
        q = fwd[found]
        r = first[pos[found]]
        ok = (ad[q] > 0) & (ad[r] > 0)
        q, r = q[ok], r[ok]
        self.conn_begin = cell[q].astype(np.int32)
        self.conn_end = nbr[q].astype(np.int32)
        self.conn_dir = dirn[q].astype(np.int32)
        self.conn_adB = ad[q].astype(np.float64)
        self.conn_adE = ad[r].astype(np.float64)
        self.nconn = len(q)


class Connections:
    """A corner-point grid's cells and BulkConn connections (bId < eId,
    sorted by (bId, eId)) with their transmissibility factors Akd: the
    interface the library's ocpg_set_conn and the deck reader take (the
    same attributes as workloads.Grid)."""

    def __init__(self, cg, kx, ky, kz, ntg=None):
        nb = cg.nb
        self.nx, self.ny, self.nz, self.nb = cg.nx, cg.ny, cg.nz, nb
        self.v = cg.v.copy()
        self.depth = cg.depth.copy()
        self.dx, self.dy, self.dz = cg.dx.copy(), cg.dy.copy(), cg.dz.copy()
        ntg = np.ones(nb) if ntg is None else np.asarray(ntg, dtype=np.float64)
        b, e, dr = cg.conn_begin.astype(np.int64), cg.conn_end.astype(np.int64), cg.conn_dir
        kb = np.choose(dr - 1, (kx[b], ky[b], kz[b]))
        ke = np.choose(dr - 1, (kx[e], ky[e], kz[e]))
        lateral = dr < 3
        # BulkConn::CalAkd (src/BulkConn.cpp:89-117): NTG on lateral faces only
        T1 = np.where(lateral, ntg[b] * kb * cg.conn_adB, kb * cg.conn_adB)
        T2 = np.where(lateral, ntg[e] * ke * cg.conn_adE, ke * cg.conn_adE)
        area = 1 / (1 / T1 + 1 / T2)
        order = np.lexsort((e, b))
        b, e = b[order], e[order]
        if len(b) > 1 and np.any((b[1:] == b[:-1]) & (e[1:] == e[:-1])):
            raise CornerError("two faces connect the same pair of cells")
        self.bId = b.astype(np.int32)
        self.eId = e.astype(np.int32)
        self.area = area[order].astype(np.float64)
        self.direction = dr[order].astype(np.int32)
        self.nconn = len(self.bId)

    def neighbour_counts(self):
        cnt = np.ones(self.nb, dtype=np.int64)
        np.add.at(cnt, self.bId, 1)
        np.add.at(cnt, self.eId, 1)
        return cnt
