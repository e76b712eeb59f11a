"""Field and marker output files: Tecplot ASCII and legacy VTK.

The reference's LBMD3.save_bak (src/funlbm/lbm/lbm3d.py:169-244; it is never
called there) writes, every file.per_steps steps,
  tecplot/tecplot_flow_{u,rho,p}_{step:010d}.dat   u, rho and p on the grid
  vtk/flow_{step:010d}                            u (points), rho and p
  vtk/particle_{i:03d}_{step:010d}                each particle's markers with lu_x
through two libraries that are not installed here (funtecplot's PointData,
funvtk's gridToVTK/pointsToVTK).  This module writes the same sets of
quantities under the same names in two published formats:
  - Tecplot ASCII, one ordered zone in POINT packing (I = x fastest):
    TITLE, VARIABLES "x", "y", "z", <fields>, ZONE I=, J=, K=, F=POINT;
  - legacy VTK (DataFile Version 3.0): STRUCTURED_POINTS with POINT_DATA for
    the flow (a vector u, scalars rho and p; the reference passes rho and p,
    which live on the grid points, as its cell data), POLYDATA vertices with
    the scalar u for the markers; BINARY (big-endian float32, x fastest) by
    default, ASCII on request.
tecplot_sparse restates the reference's sparse Tecplot writer
(src/funlbm/file/tecplot/dump.py:13-22) as it is, header and number format
included.
"""

import os

import numpy as np


def tecplot_sparse(dense, filename: str) -> str:
    """The non-zero entries of a 3-D array as rows "i j k value" under the
    reference's header (src/funlbm/file/tecplot/dump.py:4-22; the numbers
    print as numpy gives them after stacking the indices with the values)."""
    dense = np.asarray(dense)
    idx = np.nonzero(dense)
    rows = np.vstack((idx[0], idx[1], idx[2], dense[idx])).T
    shape = dense.shape
    with open(filename, "w") as f:
        f.write('TITLE = "Sparse Matrix Data"\n')
        f.write('VARIABLES = "Z", "Y", "X", "Value"\n')
        f.write(f"ZONE I={shape[2]}, J={shape[1]}, K={shape[0]}, F=POINT\n")
        for row in rows:
            f.write(f"{row[0]} {row[1]} {row[2]} {row[3]}\n")
    return filename


def _xfast(a):
    """[X, Y, Z, ...] -> rows in x-fastest order ([Z, Y, X, ...] flattened)."""
    a = np.asarray(a)
    axes = (2, 1, 0) + tuple(range(3, a.ndim))
    return np.transpose(a, axes).reshape((-1,) + a.shape[3:])


def tecplot_points(fields, names, filename: str, title: str = "flow") -> str:
    """One ordered Tecplot zone (POINT packing) on the node grid: columns
    x, y, z (node indices) and the given fields ([X, Y, Z] or [X, Y, Z, C]
    arrays, C columns each), I = x varying fastest."""
    fields = [np.asarray(f, dtype=np.float32) for f in fields]
    X, Y, Z = fields[0].shape[:3]
    cols = []
    for f in fields:
        v = _xfast(f)
        cols.append(v.reshape(len(v), -1))
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    coords = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1)
    data = np.concatenate(cols, axis=1)
    with open(filename, "w") as f:
        f.write(f'TITLE = "{title}"\n')
        f.write("VARIABLES = " + ", ".join(f'"{n}"' for n in ["x", "y", "z"] + list(names)) + "\n")
        f.write(f"ZONE I={X}, J={Y}, K={Z}, F=POINT\n")
        fmt = " ".join(["%d"] * 3 + ["%.9g"] * data.shape[1])
        np.savetxt(f, np.concatenate([coords.astype(np.float64), data.astype(np.float64)], axis=1), fmt=fmt)
    return filename


def _vtk_block(f, arr, binary):
    a = np.asarray(arr, dtype=np.float32).ravel()
    if binary:
        f.write(a.astype(">f4").tobytes())
        f.write(b"\n")
    else:
        f.write(("\n".join(" ".join(f"{v:.9g}" for v in a[i:i + 9]) for i in range(0, len(a), 9)) + "\n").encode())


def vtk_grid(filename: str, vectors: dict, scalars: dict, spacing: float = 1.0, title: str = "flow",
             binary: bool = True) -> str:
    """Legacy VTK STRUCTURED_POINTS file of point data on the node grid:
    vectors {name: [X, Y, Z, 3]}, scalars {name: [X, Y, Z]}."""
    first = next(iter(vectors.values())) if vectors else next(iter(scalars.values()))
    X, Y, Z = np.asarray(first).shape[:3]
    with open(filename, "wb") as f:
        head = [
            "# vtk DataFile Version 3.0",
            title,
            "BINARY" if binary else "ASCII",
            "DATASET STRUCTURED_POINTS",
            f"DIMENSIONS {X} {Y} {Z}",
            "ORIGIN 0 0 0",
            f"SPACING {spacing} {spacing} {spacing}",
            f"POINT_DATA {X * Y * Z}",
        ]
        f.write(("\n".join(head) + "\n").encode())
        for name, v in vectors.items():
            f.write(f"VECTORS {name} float\n".encode())
            _vtk_block(f, _xfast(v), binary)
        for name, v in scalars.items():
            f.write(f"SCALARS {name} float 1\nLOOKUP_TABLE default\n".encode())
            _vtk_block(f, _xfast(v), binary)
    return filename


def vtk_points(filename: str, points, scalars: dict, title: str = "markers", binary: bool = True) -> str:
    """Legacy VTK POLYDATA file of points (one vertex each) with point
    scalars {name: [M]}."""
    pts = np.asarray(points, dtype=np.float32).reshape(-1, 3)
    m = len(pts)
    with open(filename, "wb") as f:
        head = ["# vtk DataFile Version 3.0", title, "BINARY" if binary else "ASCII", "DATASET POLYDATA",
                f"POINTS {m} float"]
        f.write(("\n".join(head) + "\n").encode())
        _vtk_block(f, pts, binary)
        f.write(f"VERTICES {m} {2 * m}\n".encode())
        cells = np.stack([np.ones(m, np.int32), np.arange(m, dtype=np.int32)], axis=1).ravel()
        if binary:
            f.write(cells.astype(">i4").tobytes() + b"\n")
        else:
            f.write(("\n".join(f"1 {i}" for i in range(m)) + "\n").encode())
        f.write(f"POINT_DATA {m}\n".encode())
        for name, v in scalars.items():
            f.write(f"SCALARS {name} float 1\nLOOKUP_TABLE default\n".encode())
            _vtk_block(f, v, binary)
    return filename


def save_fields(lbm, step: int, binary: bool = True) -> list:
    """The file set of the reference's save_bak for one step (materialises
    u, rho and p once); returns the paths written."""
    fc = lbm.config.file_config
    tag = str(step).zfill(10)
    u = lbm.flow.u.detach().cpu().numpy()
    rho = lbm.flow.rou.detach().cpu().numpy()[..., 0]
    p = lbm.flow.p.detach().cpu().numpy()[..., 0]
    tec, vtk = fc.tecplot_path, fc.vtk_path
    out = [
        tecplot_points([u], ["ux", "uy", "uz"], os.path.join(tec, f"tecplot_flow_u_{tag}.dat"), "flow_u"),
        tecplot_points([rho], ["rho"], os.path.join(tec, f"tecplot_flow_rho_{tag}.dat"), "flow_u"),
        tecplot_points([p], ["p"], os.path.join(tec, f"tecplot_flow_p_{tag}.dat"), "flow_u"),
        vtk_grid(os.path.join(vtk, f"flow_{tag}.vtk"), {"u": u}, {"rho": rho, "p": p},
                 spacing=float(lbm.config.dx), binary=binary),
    ]
    for i, part in enumerate(lbm.particle_swarm.particles):
        lx = part.lx.detach().cpu().numpy()
        lu = part.lu.detach().cpu().numpy()
        out.append(vtk_points(os.path.join(vtk, f"particle_{str(i).zfill(3)}_{tag}.vtk"), lx, {"u": lu[:, 0]},
                              binary=binary))
    return out
