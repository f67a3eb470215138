"""Fluent case-file reader (mesh ingest for the finite-volume assembly).

Follows the reference's FluentReader (src/fvm/src/modules/importers/
FluentReader.cpp) and its Python face FluentCase: sections are read by id
(binary when id > 1000, double-precision nodes when id > 3000):
  2 dimension; 10 nodes; 12 cells; 13 faces; 39 / 45 zone names and types
  (FluentReader.cpp:423-570). Faces store their nodes then the two cells
  (1-based, 0 = outside); a boundary face gets a ghost cell numbered after the
  interior cells in the order faces are read, and its node order is reversed in
  3-D (in 2-D when its first cell is 0) so that the area vector points from c0
  to c1 (FluentReader.cpp:296-353). The mesh of a cell zone lists its interior
  face zones first, then its boundary face zones, both by ascending zone id, and
  its cells then the ghost cells of its boundary faces (FluentReader.cpp:657-780).

Meshes with one cell zone, no interface or periodic zones are supported; the
result feeds ``fvm_amd.thermal.Mesh.from_arrays`` (or the assembly oracle in
the tests) through ``FluentMesh.arrays()``.
"""
import re

import numpy as np

# Fluent boundary-condition thread types (zone types of section 13 headers)
_THREAD_TYPES = {2: "interior", 3: "wall", 4: "pressure-inlet", 5: "pressure-outlet", 7: "symmetry",
                 8: "periodic-shadow", 9: "pressure-far-field", 10: "velocity-inlet", 12: "periodic",
                 14: "fan", 20: "mass-flow-inlet", 24: "interface", 31: "parent", 36: "outflow",
                 37: "axis"}


class FluentError(ValueError):
    pass


class FluentMesh:
    """One cell zone's mesh in the reference's layout (Mesh of FluentReader::createMesh)."""

    def __init__(self, dim, nodes, face_nptr, face_nodes, face_cells, n_cells, n_ghost, group_ptr, group_ids,
                 group_types, cell_zone_id):
        self.dim = dim
        self.nodes = nodes                  # (n_nodes, 3) float64
        self.face_nptr = face_nptr          # int32 [n_faces + 1]
        self.face_nodes = face_nodes        # int32
        self.face_cells = face_cells        # int32 [2 n_faces]
        self.n_cells = n_cells
        self.n_ghost = n_ghost
        self.group_ptr = group_ptr          # int32 [n_groups + 1], group 0 interior
        self.group_ids = group_ids          # Fluent zone id of each group (0 for the interior group)
        self.group_types = group_types      # zone type names ("interior", "wall", ...)
        self.cell_zone_id = cell_zone_id

    @property
    def n_faces(self):
        return self.face_cells.size // 2

    def getDimension(self):
        return self.dim

    def getCellZoneID(self):
        return self.cell_zone_id

    def getBoundaryGroupIDs(self):
        """Zone ids of the boundary face groups (ThermalModel's BC map keys)."""
        return list(self.group_ids[1:])

    def arrays(self):
        """Host arrays in the layout of fvm_mesh_create_general / the assembly oracle."""
        return dict(dim=self.dim, nodes=self.nodes.reshape(-1).copy(), face_nptr=self.face_nptr,
                    face_nodes=self.face_nodes, face_cells=self.face_cells, n_cells=self.n_cells,
                    n_ghost=self.n_ghost, group_ptr=self.group_ptr)

    def to_device(self, device=0):
        from .thermal import Mesh
        return Mesh.from_arrays(self.nodes, self.face_nodes, self.face_cells, self.n_cells, self.n_ghost,
                                self.group_ptr, device=device, face_nptr=self.face_nptr, dim=self.dim)


class FluentCase:
    """FluentCase(fileName).read(); getMeshList() (the reference's Python reader API)."""

    def __init__(self, file_name):
        self.file_name = file_name
        self._meshes = None

    # ------------------------------------------------------------------ scanning
    def read(self):
        with open(self.file_name, "rb") as fp:
            self._d = fp.read()
        self._i = 0
        self.dim = 0
        self.n_nodes = self.n_cells = self.n_faces = 0
        self._coords = None
        self._cell_zones = {}    # id -> (iBeg, iEnd, type), 0-based inclusive
        self._face_zones = {}    # id -> dict(iBeg, iEnd, type, nodes, cells)
        self._zone_info = {}     # id -> (type name, zone name)
        while True:
            sid = self._next_section()
            if sid is None:
                break
            self._section(sid)
        self._meshes = self._build()
        return self

    def getMeshList(self):
        if self._meshes is None:
            raise FluentError("read() first")
        return self._meshes

    def _next_section(self):
        d = self._d
        n = len(d)
        while self._i < n:
            c = d[self._i]
            if c == 0x28:   # '('
                self._i += 1
                m = re.compile(rb"\s*(\d+)").match(d, self._i)
                if not m:
                    raise FluentError("bad section start at byte %d" % self._i)
                self._i = m.end()
                return int(m.group(1))
            self._i += 1
        return None

    def _skip_ws(self):
        d = self._d
        while self._i < len(d) and d[self._i] in b" \t\r\n":
            self._i += 1

    def _close_ascii(self, depth=1):
        """Skip to the ')' closing `depth` open parentheses (quoted strings skipped)."""
        d = self._d
        n = len(d)
        while self._i < n and depth > 0:
            c = d[self._i]
            if c == 0x22:   # '"'
                j = d.index(b'"', self._i + 1)
                self._i = j + 1
                continue
            if c == 0x28:
                depth += 1
            elif c == 0x29:
                depth -= 1
            self._i += 1

    def _close_binary(self, sid):
        marker = b"End of Binary Section"
        j = self._d.find(marker, self._i)
        if j < 0:
            raise FluentError("unterminated binary section %d" % sid)
        self._i = j + len(marker)
        self._close_ascii(1)

    def _header(self):
        """'(a b c ...)' of hex integers after the section id."""
        self._skip_ws()
        d = self._d
        if d[self._i] != 0x28:
            raise FluentError("section header expected at byte %d" % self._i)
        j = d.index(b")", self._i)
        vals = [int(t, 16) for t in d[self._i + 1:j].split()]
        self._i = j + 1
        return vals

    def _open_body(self):
        """Move past the '(' that opens a section body; False if the section ends first."""
        d = self._d
        while self._i < len(d):
            c = d[self._i]
            if c == 0x28:
                self._i += 1
                return True
            if c == 0x29:
                return False
            self._i += 1
        return False

    def _ascii_tokens(self, count):
        d = self._d
        out = []
        while len(out) < count:
            m = re.compile(rb"\s*([^\s()]+)").match(d, self._i)
            if not m:
                raise FluentError("unexpected end of data at byte %d" % self._i)
            out.append(m.group(1))
            self._i = m.end()
        return out

    def _ints(self, count, binary):
        if binary:
            a = np.frombuffer(self._d, dtype="<i4", count=count, offset=self._i).astype(np.int64)
            self._i += 4 * count
            return a
        return np.array([int(t, 16) for t in self._ascii_tokens(count)], np.int64)

    # ------------------------------------------------------------------ sections
    def _section(self, sid):
        binary = sid > 1000
        dp = sid > 3000
        kind = sid % 1000
        if kind == 2:
            self.dim = int(self._ascii_tokens(1)[0])
            self._close_ascii(1)
        elif kind in (39, 45):
            self._zone_name()
        elif kind == 10:
            self._nodes(binary, dp, sid)
        elif kind == 12:
            self._cells(binary, sid)
        elif kind == 13:
            self._faces(binary, sid)
        elif kind == 18:
            raise FluentError("periodic face pairs are not supported")
        elif binary:
            self._close_binary(sid)
        else:
            self._close_ascii(1)

    def _end(self, binary, sid):
        if binary:
            self._close_binary(sid)
        else:
            self._close_ascii(1)

    def _zone_name(self):
        self._skip_ws()
        d = self._d
        j = d.index(b")", self._i)
        toks = d[self._i + 1:j].split()
        self._i = j + 1
        zid = int(toks[0])
        ztype = toks[1].decode()
        zname = toks[2].decode().replace("-", "_") if len(toks) > 2 else ""
        self._zone_info[zid] = (ztype, zname)
        self._close_ascii(1)

    def _nodes(self, binary, dp, sid):
        h = self._header()
        zid, first, last = h[0], h[1], h[2]
        if zid == 0:
            self.n_nodes = last
            self._coords = np.zeros((last, 3))
            self._end(binary, sid)
            return
        nd = h[4] if len(h) > 4 else self.dim
        if self._open_body():
            cnt = last - first + 1
            if binary:
                dt = "<f8" if dp else "<f4"
                a = np.frombuffer(self._d, dtype=dt, count=cnt * nd, offset=self._i).astype(np.float64)
                self._i += a.size * (8 if dp else 4)
            else:
                # formatted coordinates are read as doubles only for DP sections
                # (the reference reads %e into float otherwise)
                a = np.array([float(t) for t in self._ascii_tokens(cnt * nd)])
                if not dp:
                    a = a.astype(np.float32).astype(np.float64)
            self._coords[first - 1:last, :nd] = a.reshape(cnt, nd)
            self._close_ascii(1)
        self._end(binary, sid)

    def _cells(self, binary, sid):
        h = self._header()
        zid, first, last, typ = h[0], h[1], h[2], h[3]
        if zid == 0:
            self.n_cells = last
        elif typ in (1, 17):
            self._cell_zones[zid] = (first - 1, last - 1, typ)
        elif typ == 32:
            self.n_cells -= last - first + 1
        else:
            raise FluentError("cell thread type %d not handled" % typ)
        self._end(binary, sid)

    def _faces(self, binary, sid):
        h = self._header()
        zid, first, last, typ = h[0], h[1], h[2], h[3]
        shape = h[4] if len(h) > 4 else -1
        if zid == 0:
            self.n_faces = last
            self._end(binary, sid)
            return
        if typ in (0, 31):
            self._end(binary, sid)
            return
        if shape < 0:
            shape = self.dim
        if not self._open_body():
            raise FluentError("face zone %d has no body" % zid)
        cnt = last - first + 1
        if shape in (0, 5):
            nodes, cells = [], []
            for _ in range(cnt):
                nn = int(self._ints(1, binary)[0])
                v = self._ints(nn + 2, binary)
                nodes.append(v[:nn])
                cells.append(v[nn:])
            nptr = np.concatenate([[0], np.cumsum([len(v) for v in nodes])])
            flat = np.concatenate(nodes) if nodes else np.zeros(0, np.int64)
            cells = np.array(cells, np.int64).reshape(cnt, 2)
        else:
            v = self._ints(cnt * (shape + 2), binary).reshape(cnt, shape + 2)
            nptr = np.arange(cnt + 1) * shape
            flat = v[:, :shape].reshape(-1)
            cells = v[:, shape:]
        self._face_zones[zid] = dict(iBeg=first - 1, iEnd=last - 1, type=typ, nptr=nptr, nodes=flat - 1,
                                     cells=cells)
        self._close_ascii(1)
        self._end(binary, sid)

    # ------------------------------------------------------------------ mesh
    def _build(self):
        if self.dim not in (2, 3):
            raise FluentError("dimension section missing")
        if len(self._cell_zones) != 1:
            raise FluentError("one cell zone is supported, found %d" % len(self._cell_zones))
        (czid, (cb, ce, _)), = self._cell_zones.items()
        nc = self.n_cells
        # faces in file order: boundary faces get ghosts numbered after the cells
        ghost = nc
        zones = {}
        for zid in sorted(self._face_zones, key=lambda z: self._face_zones[z]["iBeg"]):
            fz = self._face_zones[zid]
            c = fz["cells"]
            n = c.shape[0]
            fc = np.zeros((n, 2), np.int64)
            nptr, flat = fz["nptr"], fz["nodes"]
            counts = np.diff(nptr)
            rev = np.full(n, self.dim == 3)
            both0 = (c[:, 0] == 0) & (c[:, 1] == 0)
            if both0.any():
                raise FluentError("faces without cells are not supported")
            c0zero = c[:, 0] == 0
            rev ^= c0zero
            bnd = (c[:, 0] == 0) | (c[:, 1] == 0)
            fc[:, 0] = np.where(c0zero, c[:, 1], c[:, 0]) - 1
            fc[:, 1] = np.where(bnd, -1, c[:, 1] - 1)
            nb = int(bnd.sum())
            fc[bnd, 1] = ghost + np.arange(nb)
            ghost += nb
            out = flat.copy()
            for f in np.nonzero(rev)[0]:
                out[nptr[f]:nptr[f + 1]] = flat[nptr[f]:nptr[f + 1]][::-1]
            zones[zid] = dict(fc=fc, nptr=nptr, nodes=out, counts=counts, boundary=bool(bnd.all()),
                              interior=not bnd.any())
            if bnd.any() and not bnd.all():
                raise FluentError("face zone %d mixes boundary and interior faces" % zid)
        n_ghost_total = ghost - nc
        interior_ids = sorted(z for z in zones if zones[z]["interior"])
        boundary_ids = sorted(z for z in zones if zones[z]["boundary"])
        order = interior_ids + boundary_ids
        # cells: the zone's cells, then the ghosts of its boundary faces in face order
        gmap = np.full(nc + n_ghost_total, -1, np.int64)
        gmap[cb:ce + 1] = np.arange(ce - cb + 1)
        nloc = ce - cb + 1
        k = nloc
        for z in boundary_ids:
            g = zones[z]["fc"][:, 1]
            gmap[g] = k + np.arange(g.size)
            k += g.size
        fcells = np.concatenate([gmap[zones[z]["fc"]] for z in order]).reshape(-1)
        if (fcells < 0).any():
            raise FluentError("face cells outside the cell zone")
        counts = np.concatenate([zones[z]["counts"] for z in order])
        nptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        fnodes = np.concatenate([zones[z]["nodes"] for z in order]).astype(np.int32)
        n_int = sum(zones[z]["fc"].shape[0] for z in interior_ids)
        gptr = [0, n_int]
        for z in boundary_ids:
            gptr.append(gptr[-1] + zones[z]["fc"].shape[0])
        types = ["interior"] + [self._zone_type(z) for z in boundary_ids]
        mesh = FluentMesh(self.dim, self._coords.copy(), nptr, fnodes, fcells.astype(np.int32), nloc, k - nloc,
                          np.array(gptr, np.int32), [0] + boundary_ids, types, czid)
        return [mesh]

    def _zone_type(self, zid):
        if zid in self._zone_info:
            return self._zone_info[zid][0]
        return _THREAD_TYPES.get(self._face_zones[zid]["type"], "unknown")
