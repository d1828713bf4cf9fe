"""The field files through the C-ABI's host functions, against the
reference's own writers and reader (OutputNeutral::Write, OutputVTK::Write
and C_Restart::ReadAscii_WM1D / WM2D, built from the reference's sources by
oracle/build_ref.sh; the fixtures are tests/golden/neutral/*.dat, *.vts and
tests/golden/ref_neutral.npz, written by oracle/ref_golden.cpp):
* from the same padded fields, the mixed and the five terms' own face
  velocities and the clock, yz_write_vts writes the reference's VTK file
  byte for byte, 1-D and 2-D, without and with dimensions; the file's own
  structure (offsets, block lengths, the two layers) read back independently;
* from the same padded fields and face velocities, yz_write_dat writes the
  reference's file byte for byte, 1-D and 2-D, without and with dimensions;
* yz_read_dat puts every number of the reference's file where the
  reference's reader puts it, with its H_coef;
* a file read back gives the real cells to the bit; a file whose point
  count or rows do not fit is refused.
"""
import math
import os
import re

import numpy as np
import pytest

from yuzen_amd import capi

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
NEU = np.load(os.path.join(GOLD, "ref_neutral.npz"))


def _case(d):
    c = capi.case_defaults()
    if d == "2d":
        c.dim, c.nx, c.ny = 2, 18, 14
        c.xmin, c.xmax, c.ymin, c.ymax = -0.5, 0.5, -0.4, 0.4
    else:
        c.dim, c.nx, c.ny = 1, 28, 1
        c.xmin, c.xmax, c.ymin, c.ymax = -0.5, 0.5, -1.0, 1.0
    return c


def _fields(d):
    return [NEU[f"{d}__{k}"] for k in ("h", "s", "c", "vx", "vy")]


@pytest.mark.parametrize("d", ["1d", "2d"])
@pytest.mark.parametrize("kind", ["nd", "dim"])
def test_writer_matches_reference_byte_for_byte(d, kind, tmp_path):
    h, s, c, vx, vy = _fields(d)
    out = str(tmp_path / "f.dat")
    capi.write_dat(out, _case(d), h, s, c, vx, vy, L0=1.3e-3, H0=2.5e-6, U0=0.07, dimensioned=(kind == "dim"))
    want = open(os.path.join(GOLD, "neutral", f"{d}_{kind}.dat"), "rb").read()
    got = open(out, "rb").read()
    assert got == want
    assert not os.path.exists(out + ".tmp")


def _velocity6(d):
    ux = np.stack([NEU[f"{d}__vx"]] + [NEU[f"{d}__tx{t}"] for t in range(5)])
    uy = np.stack([NEU[f"{d}__vy"]] + [NEU[f"{d}__ty{t}"] for t in range(5)])
    return ux, uy


@pytest.mark.parametrize("d", ["1d", "2d"])
@pytest.mark.parametrize("kind", ["nd", "dim"])
def test_vts_writer_matches_reference_byte_for_byte(d, kind, tmp_path):
    h, s, c, _, _ = _fields(d)
    ux, uy = _velocity6(d)
    time, T0 = NEU[f"{d}__clock"]
    out = str(tmp_path / "f.vts")
    capi.write_vts(out, _case(d), time, h, s, c, ux, uy if d == "2d" else None, L0=1.3e-3, H0=2.5e-6, U0=0.07,
                   T0=T0, dimensioned=(kind == "dim"))
    want = open(os.path.join(GOLD, "neutral", f"{d}_{kind}.vts"), "rb").read()
    got = open(out, "rb").read()
    assert len(got) == len(want)
    assert got == want
    assert not os.path.exists(out + ".tmp")


def parse_vts(raw):
    """the header's attributes and the appended blocks of a .vts, read by
    the file's own offsets and block lengths"""
    head, _, tail = raw.partition(b'<AppendedData encoding="raw">\n   _')
    text = head.decode()
    yd = dict(re.findall(r'(\w+)="([^"]*)"', re.search(r"<YuzenData[^>]*>", text).group(0)))
    arrays = []
    for tag in re.findall(r"<DataArray [^>]*>", text):
        at = dict(re.findall(r'(\w+)="([^"]*)"', tag))
        arrays.append((at["Name"], at.get("NumberOfComponents", "1"), at["offset"]))
    blocks = {}
    for name, ncomp, off in arrays:
        off = int(off)
        n = int(np.frombuffer(tail[off:off + 4], dtype=np.uint32)[0])
        blocks[name] = np.frombuffer(tail[off + 4:off + 4 + n], dtype="<f8").reshape(-1, int(ncomp))
    end = max(int(o) for _, _, o in arrays)
    n_last = int(np.frombuffer(tail[end:end + 4], dtype=np.uint32)[0])
    assert tail[end + 4 + n_last:] == b"\n</AppendedData>\n</VTKFile>\n"
    return yd, blocks


@pytest.mark.parametrize("d", ["1d", "2d"])
def test_vts_structure(d, tmp_path):
    """the layers and the columns of the file, against the .dat of the same
    fields: every scalar repeated on both layers, Height = h + s, z = s on
    layer 0 and h + s on layer 1, the velocity blocks (u, v, 0), each term's
    block its own field's points"""
    h, s, c, vx, vy = _fields(d)
    ux, uy = _velocity6(d)
    out = str(tmp_path / "f.vts")
    capi.write_vts(out, _case(d), 0.5, h, s, c, ux, uy if d == "2d" else None)
    yd, b = parse_vts(open(out, "rb").read())
    dat = str(tmp_path / "f.dat")
    capi.write_dat(dat, _case(d), h, s, c, vx, vy)
    cols = np.loadtxt(dat, skiprows=1)
    nox = _case(d).nx + 2
    noy = _case(d).ny + 2 if d == "2d" else 1
    m = nox * noy
    assert (yd["Time"], yd["Model"], yd["Nx"], yd["Ny"], yd["Nz"]) == ("5.000000e-01", "WM2D", str(nox), str(noy), "2")
    if d == "2d":
        x, y, HS, U, V, C, H, S = cols.T
    else:
        x, HS, U, C, H, S = cols.T
        y, V = np.zeros(m), np.zeros(m)
    for name, want in (("Conc", C), ("Height", HS), ("Substrate", S)):
        assert np.array_equal(b[name][:m, 0], b[name][m:, 0])
        assert np.allclose(b[name][:m, 0], want, rtol=1e-15, atol=0)
    assert np.array_equal(b["Velocity"][:m, 0], U) and np.array_equal(b["Velocity"][:m, 1], V)
    assert np.array_equal(b["Velocity"][m:], b["Velocity"][:m]) and not b["Velocity"][:, 2].any()
    for t, name in enumerate(("V_Ca", "V_Ma", "V_Bo", "V_Pg", "V_Sp")):
        tdat = str(tmp_path / f"t{t}.dat")
        capi.write_dat(tdat, _case(d), h, s, c, ux[t + 1], uy[t + 1])
        tc = np.loadtxt(tdat, skiprows=1)
        assert np.array_equal(b[name][:m, 0], tc[:, 3 if d == "2d" else 2]), name
        if d == "2d":
            assert np.array_equal(b[name][:m, 1], tc[:, 4]), name
    P = b["Coordinate"]
    assert np.array_equal(P[:m, 0], x) and np.array_equal(P[:m, 1], y)
    assert np.array_equal(P[:m, 2], b["Substrate"][:m, 0])
    assert np.allclose(P[m:, 2], HS, rtol=1e-15, atol=0)


def test_vts_refuses_a_2d_film_without_y_velocities(tmp_path):
    h, s, c, _, _ = _fields("2d")
    ux, _ = _velocity6("2d")
    with pytest.raises((RuntimeError, ValueError)):
        capi.write_vts(str(tmp_path / "f.vts"), _case("2d"), 0.0, h, s, c, ux, None)


def _written_cells(d, n):
    nxp = int(NEU[f"{d}__nxp"])
    nyp = n // nxp
    m = np.zeros((nyp, nxp), dtype=bool)
    if d == "1d":
        m[0, 1:nxp - 1] = True
    else:
        m[1:nyp - 1, 1:nxp - 1] = True
    return m.ravel()


@pytest.mark.parametrize("d", ["1d", "2d"])
def test_reader_matches_reference(d):
    n = NEU[f"{d}__h"].size
    z = np.zeros(n)
    h, s, c = capi.read_dat(os.path.join(GOLD, "neutral", f"{d}_nd.dat"), _case(d), z, z, z, H_coef=1.7)
    m = _written_cells(d, n)
    for got, k in ((h, "h"), (s, "s"), (c, "c")):
        want = NEU[f"{d}__read_{k}"]
        assert np.array_equal(got[m], want[m]), k
        assert np.all(got[~m] == 0)


@pytest.mark.parametrize("d", ["1d", "2d"])
def test_round_trip_keeps_the_real_cells(d, tmp_path):
    h, s, c, vx, vy = _fields(d)
    out = str(tmp_path / "f.dat")
    capi.write_dat(out, _case(d), h, s, c, vx, vy)
    z = np.full(h.size, -7.0)
    h2, s2, c2 = capi.read_dat(out, _case(d), z, z, z)
    nxp = int(NEU[f"{d}__nxp"])
    nyp = h.size // nxp
    real = np.zeros((nyp, nxp), dtype=bool)
    if d == "1d":
        real[0, 2:nxp - 2] = True
    else:
        real[2:nyp - 2, 2:nxp - 2] = True
    real = real.ravel()
    for a, b in ((h, h2), (s, s2), (c, c2)):
        assert np.array_equal(a[real], b[real])


def test_bad_files_are_refused(tmp_path):
    h, s, c, vx, vy = _fields("1d")
    good = str(tmp_path / "g.dat")
    capi.write_dat(good, _case("1d"), h, s, c, vx, vy)
    lines = open(good).read().splitlines()
    z = np.zeros(h.size)
    short = tmp_path / "short.dat"
    short.write_text("\n".join(lines[:-1]) + "\n")
    with pytest.raises(RuntimeError):
        capi.read_dat(str(short), _case("1d"), z, z, z)
    bad = tmp_path / "bad.dat"
    bad.write_text("\n".join(lines[:5] + [lines[5].replace("e", "x", 1)] + lines[6:]) + "\n")
    with pytest.raises(RuntimeError):
        capi.read_dat(str(bad), _case("1d"), z, z, z)
    # a seventh column is read only under a header that names PHI
    phi = tmp_path / "phi.dat"
    phi.write_text("\n".join([lines[0] + " \t PHI[-]"] + [l + "\t1.5" for l in lines[1:]]) + "\n")
    h2, _, _ = capi.read_dat(str(phi), _case("1d"), z, z, z)
    assert np.array_equal(h2[2:-2], h[2:-2])
    nophi = tmp_path / "nophi.dat"
    nophi.write_text("\n".join([lines[0]] + [l + "\t1.5" for l in lines[1:]]) + "\n")
    with pytest.raises(RuntimeError):
        capi.read_dat(str(nophi), _case("1d"), z, z, z)


def test_equilibrium_droplet_residual_velocity(tmp_path):
    """The reference's equilibrium droplets read with yz_read_dat, their
    velocity (the oracle's, which is the reference's CalVelocity to the
    bit) written with yz_write_dat: the largest Ux is the one the
    reference's solver publishes, to its 4 digits, at 2nd order in dx."""
    from oracle import binding as ob
    from .eq_drop import PUBLISHED, case, max_ux, read
    got = {}
    for n, want in PUBLISHED.items():
        h, s, c = read(n, tmp_path)
        assert abs(h.max() - 1.0) < 1e-5
        o = ob.OracleStep(case(n), h, s, c, 1e-6)
        hg = o.h()
        vx, vy = o.velocity(hg, c)
        out = str(tmp_path / f"out{n}.dat")
        capi.write_dat(out, case(n), hg, s, c, vx, vy)
        got[n] = max_ux(out)
        print(n, got[n], want)
        assert abs(got[n] - want) <= 5e-4 * want
    assert 1.85 < math.log2(got[1000] / got[2000]) < 2.05
    assert 1.9 < math.log2(got[2000] / got[4000]) < 2.05
