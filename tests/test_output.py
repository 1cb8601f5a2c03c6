"""VTK output (charm_write_vtk / charm_write_solution, src/charm_output.c:10-126
of the reference): the piece is parsed back here with an independent reader
of the XML header and the raw appended blocks, and checked against the
leaves (shared corners, hexahedra, levels) and the oracle's cons_to_prim of
the cell-centre states (R P E E_TOT U V W T C_*).  CPU only."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import charm_3d_amd as CA
import mesh_util as MU
import oracle_lib as OL

R = CA.GAS_R


def read_vtu(path):
    raw = open(path, "rb").read()
    head, _, tail = raw.partition(b"<AppendedData encoding=\"raw\">")
    data = tail[tail.index(b"_") + 1:]
    root = ET.fromstring(head + b"</VTKFile>" if b"</VTKFile>" not in head else head)
    piece = root.find("UnstructuredGrid/Piece")
    npts, ncell = int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells"))
    dt = {"Float64": np.float64, "Int64": np.int64, "Int32": np.int32, "UInt8": np.uint8}

    def block(da):
        off = int(da.get("offset"))
        nbytes = int(np.frombuffer(data[off:off + 8], dtype=np.uint64)[0])
        return np.frombuffer(data[off + 8:off + 8 + nbytes], dtype=dt[da.get("type")])

    pts = block(piece.find("Points/DataArray")).reshape(-1, 3)
    cells = {da.get("Name"): block(da) for da in piece.find("Cells")}
    cdata = {da.get("Name"): block(da) for da in piece.find("CellData")}
    return npts, ncell, pts, cells, cdata


def write_and_read(tmp_path, leaves, u, cfg, name="sol"):
    base = str(tmp_path / name)
    CA.write_vtk(base, leaves, u, cfg)
    assert os.path.exists(base + ".pvtu")
    pv = ET.parse(base + ".pvtu").getroot()
    assert pv.find("PUnstructuredGrid/Piece").get("Source") == name + "_0000.vtu"
    return read_vtu(base + "_0000.vtu"), pv


def random_state(rng, n, nc, cfg):
    u = np.zeros((n, 4 + nc, 4))
    rho = 1 + 0.2 * rng.uniform(-1, 1, n)
    vel = rng.normal(size=(n, 3))
    p = 1 + 0.3 * rng.uniform(0, 1, n)
    y = rng.dirichlet(np.ones(nc), size=n) if nc > 1 else np.ones((n, 1))
    cp = np.array([cfg.comp[i].cp[0] for i in range(nc)])
    ms = np.array([cfg.comp[i].m for i in range(nc)])
    if cfg.eos == 0 or nc == 1:
        cpm, mm = np.full(n, cp[0]), np.full(n, ms[0])
    else:
        cpm, mm = (y * cp).sum(1), 1.0 / (y / ms).sum(1)
    gam = cpm / (cpm - R / mm)
    e = p / (rho * (gam - 1))
    u[:, 0:3, 0] = rho[:, None] * vel
    u[:, 3, 0] = rho * (e + 0.5 * (vel ** 2).sum(1))
    u[:, 4:, 0] = rho[:, None] * y
    u[:, :, 1:] = rng.normal(size=(n, 4 + nc, 3)) * 0.01
    return u


@pytest.mark.parametrize("nc,eos", [(1, "IDEAL"), (2, "MIX")])
def test_vtk_piece_matches_leaves_and_prim_fields(tmp_path, nc, eos):
    rng = np.random.default_rng(7)
    leaves = MU.random_adapted(2, 4, seed=11, frac=0.3)
    n = len(leaves)
    comp_m = (R, 2 * R)[:nc]
    comp_cp = (3.5, 2.5)[:nc]
    cfg = CA.make_config(precision=CA.FP64, eos=eos, comp_m=comp_m, comp_cp=comp_cp,
                         comp_names=("AIR", "HE")[:nc])
    u = random_state(rng, n, nc, cfg)
    (npts, ncell, pts, cells, cd), pv = write_and_read(tmp_path, leaves, u, cfg)
    assert ncell == n
    # hexahedra whose corners are the leaf corners, points shared
    assert (cells["types"] == 12).all()
    assert (cells["offsets"] == 8 * np.arange(1, n + 1)).all()
    conn = cells["connectivity"].reshape(n, 8)
    h = (MU.ROOT >> leaves[:, 3]) / MU.ROOT
    lo = leaves[:, :3] / MU.ROOT
    corners = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                        [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
    expect = lo[:, None, :] + corners[None, :, :] * h[:, None, None]
    assert np.array_equal(pts[conn], expect)
    uniq = np.unique(expect.reshape(-1, 3), axis=0)
    assert npts == len(uniq)
    assert (cd["level"] == leaves[:, 3]).all() and (cd["treeid"] == 0).all() and (cd["mpirank"] == 0).all()
    # cell-centre fields against the oracle's cons_to_prim of coefficient 0
    ocfg = OL.make_config(n_comp=nc, comp_m=comp_m, comp_cp=tuple((c,) for c in comp_cp), eos=eos)
    names = ["R", "P", "E", "E_TOT", "U", "V", "W", "T"] + ["C_" + s for s in ("AIR", "HE")[:nc]]
    assert list(cd)[3:] == names
    for i in range(0, n, max(1, n // 40)):
        p = OL.cons_to_prim(ocfg, u[i, :, 0])
        ref = [p.r, p.p, p.e, p.e_tot, p.u, p.v, p.w, p.t] + [p.c[k] for k in range(nc)]
        got = [cd[k][i] for k in names]
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-15)
    pnames = [da.get("Name") for da in pv.find("PUnstructuredGrid/PCellData")]
    assert pnames == ["treeid", "level", "mpirank"] + names


def test_vtk_polynomial_cp_temperature(tmp_path):
    """polynomial cp (Sutherland transport): T from the Newton solve,
    (cp(T) - R / M) T = e"""
    leaves = MU.uniform_leaves(2)
    n = len(leaves)
    cps = (3.3, 0.2, 0.05)
    cfg = CA.make_config(precision=CA.FP64, comp_m=(R,), comp_cp=(cps,))
    t = 1.0 + 0.5 * np.random.default_rng(3).uniform(size=n)
    rho = np.full(n, 1.2)
    e = (cps[0] + cps[1] * t + cps[2] * t * t - 1.0) * t
    u = np.zeros((n, 5, 4))
    u[:, 3, 0] = rho * e
    u[:, 4, 0] = rho
    (_, _, _, _, cd), _ = write_and_read(tmp_path, leaves, u, cfg)
    np.testing.assert_allclose(cd["T"], t, rtol=1e-9)
    cp = cps[0] + cps[1] * t + cps[2] * t * t
    np.testing.assert_allclose(cd["P"], rho * e * (cp / (cp - 1.0) - 1.0), rtol=1e-9)
    assert list(cd)[-1] == "C_0"


def test_vtk_rejects_bad_arguments(tmp_path):
    cfg = CA.make_config(precision=CA.FP64)
    with pytest.raises(CA.CharmError):
        CA.write_vtk(str(tmp_path / "no_such_dir" / "x"), MU.uniform_leaves(1), np.ones((8, 5, 4)), cfg)
