"""Visualize::PostProcess (PostProcess.h:13-93, PostProcess.cpp) for fields assembled on the device.

One visualizer per cell group writes, at every Plot(t), a VTK XML unstructured grid
`<result_dir>/<name><step>.vtu` and appends it to `<result_dir>/<name>.pvd`, the way PostProcess
does. Geometry and data follow AverageHandler (AverageHandler.cpp): every cell gets its own corner
points, dof values are written at the corners (point data; 2-component vectors padded to 3 so that
"warp by vector" works), and integration-point fields (Cell::Eval, e.g. Context.eval_ip of damage or
kappa) become cell data, the mean over the cell's integration points. 6-component Voigt data is
reordered from [xx, yy, zz, yz, zx, xy] to VTK's [xx, yy, zz, xy, yz, xz] and every array is
Float32, as XMLWriter.cpp writes them. Binary output is VTK's appended raw format with UInt64
block headers (XMLWriter.cpp:137-200); ASCII writes the values inline.

Host code: device tensors are copied to the host at Plot time.
"""
import os

import numpy as np

from .binding import HEX8, QUAD4, TET4

# VTK cell types and corner counts of the coordinate elements (XMLWriter.cpp:8-40)
VTK_CELL = {QUAD4: (9, 4), HEX8: (12, 8), TET4: (10, 4)}


def _host(a):
    if a is None:
        return None
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a)


def _vtk_voigt(v):
    """NuTo's Voigt order [xx, yy, zz, yz, zx, xy] -> VTK's [xx, yy, zz, xy, yz, xz]."""
    if v.shape[-1] != 6:
        return v
    return v[..., [0, 1, 2, 5, 3, 4]]


class _Visualizer:
    def __init__(self, coord_type, coords, conn):
        if coord_type not in VTK_CELL:
            raise ValueError("coordinate element without a VTK cell type")
        self.vtk_type, self.nc = VTK_CELL[coord_type]
        X = np.asarray(coords, np.float64)
        C = np.asarray(conn, np.int64)
        if C.ndim != 2 or C.shape[1] < self.nc:
            raise ValueError("connectivity does not fit the coordinate element")
        pts = X[C[:, :self.nc]].reshape(-1, X.shape[1])
        self.points = np.zeros((pts.shape[0], 3))
        self.points[:, :X.shape[1]] = pts
        self.n_cells = C.shape[0]
        self.dofs = []       # (name, values, cell_dofs, ncomp)
        self.cell_fns = []   # (name, callable -> (n_cells, n_ip[, n_comp]))


class PostProcess:
    def __init__(self, result_dir):
        """PostProcess(resultDir): creates the directory."""
        self.result_dir = ""
        self._vis = {}
        self._pvd = {}
        self._step = 0
        self.set_result_directory(result_dir)

    def set_result_directory(self, result_dir):
        self.result_dir = str(result_dir)
        if self.result_dir:
            os.makedirs(self.result_dir, exist_ok=True)

    def result_directory(self):
        return self.result_dir

    def define_visualizer(self, name, coord_type, coords, conn):
        """DefineVisualizer(name, cells, AverageHandler()): the cell group's coordinate elements."""
        if name in self._vis:
            raise ValueError(f"visualizer {name} already defined")
        self._vis[name] = _Visualizer(coord_type, coords, conn)

    def _get(self, name):
        if name not in self._vis:
            raise KeyError(f"There is no visualizer named {name}")
        return self._vis[name]

    def add_dof(self, name, dof_name, values, cell_dofs, ncomp):
        """Add(name, dof): point data from a dof vector (host array or device tensor, read at each
        Plot) through the cell dof numbering (n_cells x nodes * ncomp, node-major)."""
        v = self._get(name)
        cd = np.asarray(cell_dofs, np.int64)
        if cd.shape[0] != v.n_cells or cd.shape[1] < v.nc * ncomp:
            raise ValueError("cell dof numbering does not fit the visualizer's cells")
        v.dofs.append((dof_name, values, cd, ncomp))

    def add_cell_function(self, name, function, field_name):
        """Add(name, cellFunction, name): function() returns the field at every (cell, IP),
        shape (n_cells, n_ip) or (n_cells, n_ip, n_comp), e.g. lambda: ctx.eval_ip(IP_DAMAGE, u, e, k)."""
        self._get(name).cell_fns.append((field_name, function))

    def plot(self, t, as_binary=True):
        """Plot(t, asBinary): one .vtu per visualizer, appended to its .pvd."""
        if not self.result_dir:
            raise RuntimeError("You have to set the result directory by calling ResultDirectory(name) with a "
                               "non-empty name.")
        for name, v in self._vis.items():
            point_data = []
            for dof_name, values, cd, ncomp in v.dofs:
                vals = _host(values)
                idx = cd[:, :v.nc * ncomp].reshape(v.n_cells * v.nc, ncomp)
                pd = vals[idx]
                if ncomp == 2:
                    pd = np.concatenate([pd, np.zeros((pd.shape[0], 1))], axis=1)
                point_data.append((dof_name, pd))
            cell_data = []
            for field_name, fn in v.cell_fns:
                f = _host(fn())
                if f.ndim == 2:
                    f = f[:, :, None]
                if f.shape[0] != v.n_cells:
                    raise ValueError(f"{field_name}: expected {v.n_cells} cells, got {f.shape[0]}")
                cell_data.append((field_name, _vtk_voigt(f.mean(axis=1))))
            vtu = f"{name}{self._step}.vtu"
            write_vtu(os.path.join(self.result_dir, vtu), v.points, v.vtk_type, v.nc, point_data, cell_data,
                      as_binary)
            self._pvd.setdefault(name, []).append((t, vtu))
            write_pvd(os.path.join(self.result_dir, name + ".pvd"), self._pvd[name])
        self._step += 1


def _arrays(points, vtk_type, nc, point_data, cell_data):
    n_cells = points.shape[0] // nc
    conn = np.arange(points.shape[0], dtype=np.uint32)
    offsets = (np.arange(1, n_cells + 1, dtype=np.uint64) * nc).astype(np.uint32)
    types = np.full(n_cells, vtk_type, np.uint8)
    pd = [(n, np.asarray(a, np.float32).reshape(a.shape[0], -1)) for n, a in point_data]
    cd = [(n, np.asarray(a, np.float32).reshape(a.shape[0], -1)) for n, a in cell_data]
    return n_cells, conn, offsets, types, pd, cd, np.asarray(points, np.float32)


_VTK_TYPE = {np.dtype(np.float32): "Float32", np.dtype(np.uint32): "UInt32", np.dtype(np.uint8): "UInt8"}


def write_vtu(path, points, vtk_type, nc, point_data, cell_data, as_binary=True):
    """An UnstructuredGrid .vtu of cells with their own corner points (XMLWriter::Export)."""
    n_cells, conn, offsets, types, pd, cd, pts = _arrays(points, vtk_type, nc, point_data, cell_data)
    blocks = []

    def array(name, a, ncomp):
        t = _VTK_TYPE[a.dtype]
        nc_attr = f' NumberOfComponents="{ncomp}"' if ncomp else ""
        if as_binary:
            off = sum(8 + b.nbytes for b in blocks)
            blocks.append(np.ascontiguousarray(a).ravel())
            return f'    <DataArray type="{t}" Name="{name}"{nc_attr} format="appended" offset="{off}"/>\n'
        body = " ".join(repr(float(x)) if t == "Float32" else str(int(x)) for x in a.ravel())
        return f'    <DataArray type="{t}" Name="{name}"{nc_attr} format="ascii">\n{body}\n    </DataArray>\n'

    s = ['<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian"'
         + (' header_type="UInt64"' if as_binary else "") + '>\n <UnstructuredGrid>\n',
         f'  <Piece NumberOfPoints="{pts.shape[0]}" NumberOfCells="{n_cells}">\n', '   <PointData>\n']
    s += [array(n, a, a.shape[1]) for n, a in pd]
    s += ['   </PointData>\n', '   <CellData>\n']
    s += [array(n, a, a.shape[1]) for n, a in cd]
    s += ['   </CellData>\n', '   <Points>\n', array("points", pts, 3), '   </Points>\n', '   <Cells>\n',
          array("connectivity", conn, 0), array("offsets", offsets, 0), array("types", types, 0), '   </Cells>\n',
          '  </Piece>\n </UnstructuredGrid>\n']
    with open(path, "wb") as f:
        f.write("".join(s).encode())
        if as_binary:
            f.write(b' <AppendedData encoding="raw">\n_')
            for b in blocks:
                f.write(np.uint64(b.nbytes).tobytes())
                f.write(b.astype(b.dtype.newbyteorder("<"), copy=False).tobytes())
            f.write(b'\n </AppendedData>\n')
        f.write(b'</VTKFile>\n')


def write_pvd(path, entries):
    """The .pvd collection of (time, vtu file) entries (PostProcess::AddToPvdFile)."""
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1">\n<Collection>\n')
        for t, vtu in entries:
            f.write(f'<DataSet timestep="{t!r}" file="{vtu}"/>\n')
        f.write('</Collection>\n</VTKFile>\n')
