"""Host-side logic that runs without a GPU: mesh tables, file formats, the
icosphere and synthetic meshes, parameter schedules and the no-fallback rules.

These are the CPU halves of the reference's own tests
(tests/test_surface_stencil.py, tests/test_rotation_feature.py,
tests/test_dartel_grid.py, tests/test_surface_rotation.py): every table that
is uploaded to the device is checked here against a brute-force construction.
"""

from __future__ import annotations

import re
import numpy as np
import pytest
import torch

import cat_surface_gpu as csg
from cat_surface_gpu import gifti
from cat_surface_gpu.dartel_grid import DeviceUnavailable, resolve_device
from cat_surface_gpu.rotation_pipeline import (read_rotation_points, read_rotation_values,
                                               write_rotation_values)
from cat_surface_gpu.stencil_builder import icosphere, subdivisions_for, uv_points
from cat_surface_gpu.surface_stencil import (SurfaceStencil, build_incident_csr,
                                             build_neighbour_csr, greedy_colours, topology_for)
from cat_surface_gpu.surface_warp import default_dartel_parameters
from cat_surface_gpu.synthetic import fibonacci_hull

CPU = torch.device("cpu")


def _edges(faces):
    e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
    return np.unique(np.sort(e, axis=1), axis=0)


@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_icosphere_counts_and_orientation(k):
    v, f = icosphere(k)
    assert v.shape == (10 * 4 ** k + 2, 3) and f.shape == (20 * 4 ** k, 3)
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-15)
    e = _edges(f)
    assert v.shape[0] - e.shape[0] + f.shape[0] == 2  # Euler characteristic of a sphere
    tri = v[f]
    normal = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    assert (np.einsum("ij,ij->i", normal, tri.mean(axis=1)) > 0).all()  # outward winding
    assert subdivisions_for(f.shape[0]) == k


def test_subdivisions_for_rejects_non_icosphere_counts():
    with pytest.raises(ValueError):
        subdivisions_for(81921)


def test_uv_points_follow_pixel_centres():
    p = uv_points(8, 4).reshape(4, 8, 3)
    np.testing.assert_allclose(np.linalg.norm(p, axis=-1), 1.0, atol=1e-15)
    # row y has polar angle π(y+½)/ny; column x has azimuth 2π(x+½)/nx
    np.testing.assert_allclose(p[:, 0, 2], np.cos(np.pi * (np.arange(4) + 0.5) / 4))
    az = np.arctan2(p[1, :, 1], p[1, :, 0]) % (2 * np.pi)
    np.testing.assert_allclose(az, 2 * np.pi * (np.arange(8) + 0.5) / 8)


def test_neighbour_and_incident_tables_match_brute_force():
    v, f = icosphere(2)
    n = v.shape[0]
    off, nbr = build_neighbour_csr(f, n)
    e = _edges(f)
    want = [set() for _ in range(n)]
    for a, b in e:
        want[a].add(b)
        want[b].add(a)
    for i in range(n):
        assert set(nbr[off[i]:off[i + 1]].tolist()) == want[i]
        assert len(nbr[off[i]:off[i + 1]]) == len(want[i])
    ioff, inc = build_incident_csr(f, n)
    seen = np.zeros(3 * f.shape[0], np.int64)
    for i in range(n):
        for entry in inc[ioff[i]:ioff[i + 1]]:
            face, corner = divmod(int(entry), 3)
            assert f[face, corner] == i
            seen[entry] += 1
    assert (seen == 1).all()


def test_greedy_colouring_is_proper_and_matches_reference_fixture(golden):
    g = golden("rotation_feature")
    faces = g["faces"].astype(np.int32)
    n = g["verts"].shape[0]
    off, nbr = build_neighbour_csr(faces, n)
    colours = greedy_colours(off, nbr)
    np.testing.assert_array_equal(colours, g["colours"])
    for a, b in _edges(faces):
        assert colours[a] != colours[b]


def test_depth_system_merge_equals_dense_assembly():
    """Merged CSR entries sum exactly the COO contributions that share (row, col)."""

    v, f = icosphere(2)
    n, nf = v.shape[0], f.shape[0]
    topo = topology_for(f, n, CPU)
    assert topology_for(f, n, CPU) is topo  # cached per mesh
    rng = np.random.default_rng(0)
    coo = rng.normal(size=9 * nf)
    area = rng.uniform(0.5, 1.5, size=n)
    i0, i1, i2 = (f[:, k].astype(np.int64) for k in range(3))
    rows = np.concatenate([i0, i1, i2, i0, i1, i1, i2, i0, i2, np.arange(n)])
    cols = np.concatenate([i0, i1, i2, i1, i0, i2, i1, i2, i0, np.arange(n)])
    vals = np.concatenate([coo, 1e-3 * area])
    dense = np.zeros((n, n))
    np.add.at(dense, (rows, cols), vals)
    row_off = topo.sys_row_off.numpy()
    mcols = topo.sys_cols.numpy()
    coff = topo.sys_contrib_off.numpy()
    contrib = topo.sys_contrib.numpy()
    assert sorted(contrib.tolist()) == list(range(rows.size))
    merged = np.zeros((n, n))
    for i in range(n):
        cs = mcols[row_off[i]:row_off[i + 1]]
        assert (np.diff(cs) > 0).all()  # sorted, no duplicate columns
        for m in range(row_off[i], row_off[i + 1]):
            ks = contrib[coff[m]:coff[m + 1]]
            assert (rows[ks] == i).all() and (cols[ks] == mcols[m]).all()
            assert (np.diff(ks) > 0).all()  # contributions kept in COO order
            merged[i, mcols[m]] = vals[ks].sum()
    np.testing.assert_allclose(merged, dense, rtol=0, atol=1e-14)
    # colour groups partition the vertices
    cverts = topo.colour_verts.numpy()
    assert sorted(cverts.tolist()) == list(range(n))
    assert topo.colour_off[0] == 0 and topo.colour_off[-1] == n
    # the colour-ordered ELL holds each row's merged entries in CSR order
    w = topo.ell_width
    ecols = topo.ell_cols.numpy().reshape(w, n)
    esrc = topo.ell_src.numpy().reshape(w, n)
    for r, i in enumerate(cverts):
        want = list(range(row_off[i], row_off[i + 1]))
        got = [int(e) for e in esrc[:, r] if e >= 0]
        assert got == want
        assert (esrc[len(want):, r] == -1).all()
        np.testing.assert_array_equal(ecols[:len(want), r], mcols[want])
        assert (ecols[len(want):, r] == i).all()


def test_stencil_file_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    v, f = icosphere(1)
    st = SurfaceStencil(
        sphere_points=v, faces=f, surface_indices=rng.integers(0, 50, (v.shape[0], 3),
                                                               dtype=np.int32),
        surface_weights=rng.random((v.shape[0], 3)),
        sheet_indices=rng.integers(0, f.shape[0], (4, 8, 3), dtype=np.int32),
        sheet_weights=rng.random((4, 8, 3)), nx=8, ny=4, source_points=50,
        unit_sphere_points=rng.random((50, 3)).astype(np.float32))
    path = tmp_path / "s.stencil"
    st.to_file(path)
    back = SurfaceStencil.from_file(path)
    for name in ("sphere_points", "faces", "surface_indices", "surface_weights",
                 "sheet_indices", "sheet_weights", "unit_sphere_points"):
        np.testing.assert_array_equal(getattr(back, name), getattr(st, name))
    assert (back.nx, back.ny, back.source_points) == (8, 4, 50)
    raw = path.read_bytes()
    (tmp_path / "short").write_bytes(raw[:10])
    with pytest.raises(ValueError):
        SurfaceStencil.from_file(tmp_path / "short")
    (tmp_path / "magic").write_bytes(b"\0\0\0\0" + raw[4:])
    with pytest.raises(ValueError):
        SurfaceStencil.from_file(tmp_path / "magic")


def test_gifti_round_trip(tmp_path):
    v, f = icosphere(1)
    path = tmp_path / "m.gii"
    gifti.write_gifti_surface(path, v.astype(np.float32), f)
    arrays = gifti.read_gifti(path)
    assert [a.intent for a in arrays] == [gifti.INTENT_POINTSET, gifti.INTENT_TRIANGLE]
    np.testing.assert_array_equal(arrays[0].data, v.astype(np.float32))
    np.testing.assert_array_equal(arrays[1].data, f)
    (tmp_path / "bad.gii").write_text("<NOTGIFTI/>")
    with pytest.raises(ValueError):
        gifti.read_gifti(tmp_path / "bad.gii")


_GIFTI_ARRAY = (' <DataArray Intent="{intent}" DataType="{dtype}" ArrayIndexingOrder="{order}"'
                ' Dimensionality="2" Dim0="{d0}" Dim1="{d1}" Encoding="{enc}" Endian="{endian}"'
                ' ExternalFileName="" ExternalFileOffset="">{meta}\n  <Data>{data}</Data>\n'
                ' </DataArray>\n')


def _gifti_text(arrays, head="", root_extra="", tail=""):
    import base64
    import zlib
    parts = []
    for arr, intent, enc, endian, order, meta in arrays:
        dtype = {"f": "NIFTI_TYPE_FLOAT32", "i": "NIFTI_TYPE_INT32"}[arr.dtype.kind]
        raw = np.asarray(arr, arr.dtype.newbyteorder("<" if endian == "LittleEndian" else ">"))
        raw = raw.tobytes(order="F" if order == "ColumnMajorOrder" else "C")
        if enc == "ASCII":
            flat = arr.flatten(order="F" if order == "ColumnMajorOrder" else "C")
            data = " ".join(repr(x) for x in flat.tolist())
        elif enc == "Base64Binary":
            data = base64.b64encode(raw).decode()
        else:
            data = base64.b64encode(zlib.compress(raw, 6)).decode()
        parts.append(_GIFTI_ARRAY.format(intent=intent, dtype=dtype, order=order, d0=arr.shape[0],
                                         d1=arr.shape[1], enc=enc, endian=endian, meta=meta,
                                         data=data))
    return ('<?xml version="1.0" encoding="UTF-8"?>\n' + head +
            f'<GIFTI Version="1.0" NumberOfDataArrays="{len(arrays)}"{root_extra}>\n'
            ' <MetaData>\n  <MD>\n   <Name><![CDATA[AnatomicalStructurePrimary]]></Name>\n'
            '   <Value><![CDATA[CortexLeft]]></Value>\n  </MD>\n </MetaData>\n <LabelTable/>\n'
            + "".join(parts) + tail + "</GIFTI>\n")


def test_gifti_base64_decoder_equals_binascii():
    """The numpy base64 decoder and encoder (which release the interpreter lock)
    give the bytes of base64.b64decode / b64encode for every length and padding;
    the decoder hands text with anything but alphabet characters and end
    padding to base64.b64decode."""

    import base64

    rng = np.random.default_rng(9)
    for n in list(range(0, 24)) + [3069, 3070, 3071, 3072, 3073, 3074, 12288, 100001]:
        data = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        text = base64.b64encode(data)
        assert gifti._b64decode(text) == data
        assert gifti._b64decode(text.decode("ascii")) == data
        for odd in (text[:len(text) // 2] + b"\n" + text[len(text) // 2:], text + b"\n"):
            assert gifti._b64decode(odd) == base64.b64decode(odd)
        assert gifti._b64encode(data) == text.decode("ascii")


def test_gifti_byte_scan_equals_the_xml_parser(tmp_path):
    """Files of the plain shape are read by byte search instead of an XML parser:
    every attribute and Data text equals xml.etree's, for each encoding, byte
    order, index order and metadata form (CDATA names, a coordinate transform
    with DataSpace children, a self-closing Data); files outside that shape
    (comments, references, CDATA or markup in a payload) go through the parser
    and read the same arrays."""

    rng = np.random.default_rng(5)
    v = rng.standard_normal((37, 3)).astype(np.float32)
    f = rng.integers(0, 37, (23, 3)).astype(np.int32)
    xform = ('\n  <CoordinateSystemTransformMatrix>\n   <DataSpace><![CDATA[NIFTI_XFORM_UNKNOWN]]>'
             '</DataSpace>\n   <TransformedSpace><![CDATA[NIFTI_XFORM_UNKNOWN]]></TransformedSpace>'
             '\n   <MatrixData>1 0 0 0 0 1 0 0 0 0 1 0 0 0 0 1</MatrixData>\n'
             '  </CoordinateSystemTransformMatrix>')
    plain = []
    for enc in ("GZipBase64Binary", "Base64Binary", "ASCII"):
        for endian in ("LittleEndian", "BigEndian"):
            for order in ("RowMajorOrder", "ColumnMajorOrder"):
                plain.append(_gifti_text(
                    [(v, "NIFTI_INTENT_POINTSET", enc, endian, order, xform),
                     (f, "NIFTI_INTENT_TRIANGLE", enc, endian, order, "\n  <MetaData/>")],
                    head='<!DOCTYPE GIFTI SYSTEM "http://www.nitrc.org/frs/download.php/115/gifti.dtd">\n'))
    one = [(v, "NIFTI_INTENT_POINTSET", "GZipBase64Binary", "LittleEndian", "RowMajorOrder", "")]
    empty = _gifti_text([(np.zeros((0, 3), np.float32), "NIFTI_INTENT_POINTSET", "ASCII",
                          "LittleEndian", "RowMajorOrder", "")])
    plain.append(re.sub(r"<Data>\s*</Data>", "<Data/>", empty))
    parser_only = [
        _gifti_text(one, tail=" <!-- a <DataArray> in a comment -->\n"),
        _gifti_text(one).replace('Endian="LittleEndian"', 'Endian="Little&#69;ndian"'.replace(
            "&#69;", "E") + ' Note="a &amp; b"', 1),
        _gifti_text(one).replace("<Data>", "<Data><![CDATA[", 1).replace("</Data>", "]]></Data>", 1),
        _gifti_text(one, head='<!DOCTYPE GIFTI [ <!ENTITY e "LittleEndian"> ]>\n').replace(
            'Endian="LittleEndian"', 'Endian="&e;"', 1),
        _gifti_text([(v, "NIFTI_INTENT_POINTSET", "ASCII", "LittleEndian", "RowMajorOrder", "")]
                    ).replace("<Data>", "<Data>&#32;", 1),
    ]
    for k, text in enumerate(plain + parser_only):
        path = tmp_path / f"g{k}.gii"
        path.write_text(text)
        raw = path.read_bytes()
        scanned = gifti._scan(raw)
        parsed = gifti._elements_xml(raw, path)
        if k < len(plain):
            assert scanned is not None, k
            assert [a for a, _ in scanned] == [a for a, _ in parsed]
            assert [t.strip().decode() for _, t in scanned] == [t.strip() for _, t in parsed]
        else:
            assert scanned is None, k
        gifti.clear_caches()
        arrays = gifti.read_gifti(path)
        gifti.clear_caches()
        expected = [gifti._decode(item) for item in parsed]
        assert [a.intent for a in arrays] == [a.intent for a in expected]
        for a, b in zip(arrays, expected):
            np.testing.assert_array_equal(a.data, b.data)
        if k < len(plain) - 1 or k == len(plain) + 1:
            np.testing.assert_array_equal(arrays[0].data, v)


def test_rotation_value_and_point_files(tmp_path):
    rng = np.random.default_rng(2)
    s, t = rng.normal(size=7), rng.normal(size=5)
    write_rotation_values(tmp_path / "v.bin", s, t)
    s2, t2 = read_rotation_values(tmp_path / "v.bin")
    np.testing.assert_array_equal(s2, s)
    np.testing.assert_array_equal(t2, t)
    with (tmp_path / "v.bin").open("ab") as stream:
        stream.write(b"x")
    with pytest.raises(ValueError):
        read_rotation_values(tmp_path / "v.bin")
    pts = rng.normal(size=(6, 3))
    for dtype in ("<f4", "<f8"):
        with (tmp_path / "p.bin").open("wb") as stream:
            np.asarray([6], "<i4").tofile(stream)
            pts.astype(dtype).tofile(stream)
        np.testing.assert_array_equal(read_rotation_points(tmp_path / "p.bin"),
                                      pts.astype(dtype))
    with (tmp_path / "p.bin").open("wb") as stream:
        np.asarray([6], "<i4").tofile(stream)
        stream.write(b"\0" * 13)
    with pytest.raises(ValueError):
        read_rotation_points(tmp_path / "p.bin")


def test_default_parameter_schedule():
    """mu = 0.125, divided by 1.25 after every 4th loop; lambda divided by 5 each loop."""

    p = default_dartel_parameters(6, lambda_=1.0)
    assert [q[2] for q in p] == [0.125] * 4 + [0.1] * 2
    np.testing.assert_allclose([q[3] for q in p], [5.0 ** -i for i in range(6)])
    np.testing.assert_allclose([q[4] for q in p], [0.5 * 5.0 ** -i for i in range(6)])
    assert all(q[:2] == (1.0, 1.0) for q in p)
    with pytest.raises(ValueError):
        default_dartel_parameters(0)


def test_no_cpu_fallback():
    with pytest.raises(DeviceUnavailable):
        resolve_device("cpu")
    with pytest.raises(DeviceUnavailable):
        csg.composition(torch.zeros(2, 4, 4, dtype=torch.float64),
                        torch.zeros(2, 4, 4, dtype=torch.float64), device="cpu")
    from cat_surface_gpu.dartel_grid import _check_kernel

    _check_kernel("hip")
    _check_kernel("auto")
    for name in ("torch", "triton", "bogus"):
        with pytest.raises(ValueError):
            _check_kernel(name)


@pytest.mark.parametrize("seed", [0, 1])
def test_fibonacci_hull_is_a_closed_outward_mesh(seed):
    v, f = fibonacci_hull(642, 0.2, seed)
    assert v.shape == (642, 3)
    e = _edges(f)
    assert v.shape[0] - e.shape[0] + f.shape[0] == 2
    tri = v[f]
    normal = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    assert (np.einsum("ij,ij->i", normal, tri.mean(axis=1)) > 0).all()


def test_euler_angles_invert_the_reference_rotation_matrix():
    """angles_from_matrix inverts the oracle's Rz(Ry Rx) column-major layout."""

    from oracle import oracle as O
    from cat_surface_gpu.surface_rotation import angles_from_matrix

    rng = np.random.default_rng(11)
    for a in rng.uniform(-1.5, 1.5, size=(50, 3)):
        m = O.rotation_matrix_from_angles(a)
        b = angles_from_matrix(m)
        np.testing.assert_allclose(O.rotation_matrix_from_angles(b), m, atol=1e-13)
        np.testing.assert_allclose(b, a, atol=1e-12)   # |beta| < pi/2: the same triple


def test_hopf_grid_is_a_uniform_rotation_grid():
    """8,192 proper rotations, all distinct, with the spacing of a uniform grid."""

    from cat_surface_gpu.surface_rotation import hopf_grid, quaternion_matrix

    g = hopf_grid()
    assert g.shape == (8192, 3, 3)
    np.testing.assert_allclose(np.einsum("nij,nkj->nik", g, g), np.broadcast_to(np.eye(3),
                                                                                 g.shape),
                               atol=1e-12)
    np.testing.assert_allclose(np.linalg.det(g), 1.0, atol=1e-12)
    # geodesic distance between rotations: arccos((trace(A Bᵀ) - 1) / 2)
    tr = np.einsum("aij,bij->ab", g[:256], g)
    dist = np.arccos(np.clip((tr - 1.0) / 2.0, -1.0, 1.0))
    dist[np.arange(256), np.arange(256)] = np.inf
    nearest = dist.min(axis=1)
    assert nearest.min() > 0.15          # no duplicates (q and -q are one rotation)
    assert nearest.max() < 0.25          # ~ (8 pi^2 / 8192)^(1/3) spacing
    # covering: random rotations lie within ~0.17 rad of a grid rotation; a
    # uniform grid of 8192 covers SO(3) at about that radius
    rng = np.random.default_rng(2)
    r = quaternion_matrix(rng.normal(size=(500, 4)))
    cover = np.arccos(np.clip((np.einsum("aij,bij->ab", r, g) - 1.0) / 2.0, -1, 1)).min(axis=1)
    assert cover.max() < 0.2


def test_axis_angle_matrix_is_the_matrix_exponential():
    import scipy.linalg

    from cat_surface_gpu.surface_rotation import axis_angle_matrix

    for d in (np.array([0.3, -0.2, 0.5]), np.array([1e-9, 0.0, 2e-9]), np.zeros(3)):
        k = np.array([[0.0, -d[2], d[1]], [d[2], 0.0, -d[0]], [-d[1], d[0], 0.0]])
        np.testing.assert_allclose(axis_angle_matrix(d), scipy.linalg.expm(k), atol=1e-14)


def test_real_sh_basis_matches_scipy_spherical_harmonics():
    import math

    from scipy.special import sph_harm_y

    from cat_surface_gpu.synthetic import real_sh_basis

    rng = np.random.default_rng(4)
    p = rng.normal(size=(300, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    theta, phi = np.arccos(np.clip(p[:, 2], -1, 1)), np.arctan2(p[:, 1], p[:, 0])
    cols = []
    for l in range(1, 9):
        for m in range(-l, l + 1):
            y = sph_harm_y(l, abs(m), theta, phi)
            cols.append(math.sqrt(2) * (-1) ** m * y.imag if m < 0 else
                        (y.real if m == 0 else math.sqrt(2) * (-1) ** m * y.real))
    np.testing.assert_allclose(real_sh_basis(p, 8), np.stack(cols, axis=1), atol=1e-13)
