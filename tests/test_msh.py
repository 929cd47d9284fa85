"""Gmsh .msh input in every on-disk variant the reader accepts.

The binary files here are encoded by the test itself from the ASCII text of a
mesh, field by field as the Gmsh file format defines them (format 4.1: size_t
counts and tags, int entity fields, double coordinates; format 2.2: int
records), in both byte orders; the reader must return the same mesh as from
the ASCII text. The product writer is checked by write -> read round trips.
"""

import os

import numpy as np
import pytest

import bem_gmsh_amd as B

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _encode_41_binary(text: str, endian: str) -> bytes:
    """ASCII format 4.1 -> binary format 4.1 (data size 8)."""
    L = text.split("\n")
    out = []
    i32, u64, f64 = endian + "i4", endian + "u8", endian + "f8"
    pk = lambda dt, vals: np.array(vals, dtype=dt).tobytes()
    i = 0
    while i < len(L):
        t = L[i].strip()
        if t == "$MeshFormat":
            out += [b"$MeshFormat\n4.1 1 8\n", pk(i32, [1]), b"\n$EndMeshFormat\n"]
            i += 3
        elif t == "$PhysicalNames":
            cnt = int(L[i + 1])
            out.append(("\n".join(L[i:i + 2 + cnt + 1]) + "\n").encode())
            i += 3 + cnt
        elif t == "$Entities":
            hdr = [int(v) for v in L[i + 1].split()]
            out += [b"$Entities\n", pk(u64, hdr)]
            r = i + 2
            for _ in range(hdr[0]):
                p = L[r].split()
                nph = int(p[4])
                out += [pk(i32, [int(p[0])]), pk(f64, [float(v) for v in p[1:4]]), pk(u64, [nph]),
                        pk(i32, [int(v) for v in p[5:5 + nph]])]
                r += 1
            for _ in range(hdr[1] + hdr[2] + hdr[3]):
                p = L[r].split()
                nph = int(p[7])
                nb = int(p[8 + nph])
                out += [pk(i32, [int(p[0])]), pk(f64, [float(v) for v in p[1:7]]), pk(u64, [nph]),
                        pk(i32, [int(v) for v in p[8:8 + nph]]), pk(u64, [nb]),
                        pk(i32, [int(v) for v in p[9 + nph:9 + nph + nb]])]
                r += 1
            out.append(b"\n$EndEntities\n")
            i = r + 1
        elif t == "$Nodes":
            hdr = [int(v) for v in L[i + 1].split()]
            out += [b"$Nodes\n", pk(u64, hdr)]
            r = i + 2
            for _ in range(hdr[0]):
                d, tg, par, cnt = [int(v) for v in L[r].split()]
                out += [pk(i32, [d, tg, par]), pk(u64, [cnt])]
                out.append(pk(u64, [int(L[r + 1 + q]) for q in range(cnt)]))
                for q in range(cnt):
                    out.append(pk(f64, [float(v) for v in L[r + 1 + cnt + q].split()]))
                r += 1 + 2 * cnt
            out.append(b"\n$EndNodes\n")
            i = r + 1
        elif t == "$Elements":
            hdr = [int(v) for v in L[i + 1].split()]
            out += [b"$Elements\n", pk(u64, hdr)]
            r = i + 2
            for _ in range(hdr[0]):
                d, tg, et, cnt = [int(v) for v in L[r].split()]
                out += [pk(i32, [d, tg, et]), pk(u64, [cnt])]
                for q in range(cnt):
                    out.append(pk(u64, [int(v) for v in L[r + 1 + q].split()]))
                r += 1 + cnt
            out.append(b"\n$EndElements\n")
            i = r + 1
        else:
            i += 1
    return b"".join(out)


def _encode_22_binary(text: str, endian: str) -> bytes:
    """ASCII format 2.2 -> binary format 2.2 (one element block per record)."""
    L = text.split("\n")
    out = []
    i32, f64 = endian + "i4", endian + "f8"
    pk = lambda dt, vals: np.array(vals, dtype=dt).tobytes()
    i = 0
    while i < len(L):
        t = L[i].strip()
        if t == "$MeshFormat":
            out += [b"$MeshFormat\n2.2 1 8\n", pk(i32, [1]), b"\n$EndMeshFormat\n"]
            i += 3
        elif t == "$PhysicalNames":
            cnt = int(L[i + 1])
            out.append(("\n".join(L[i:i + 2 + cnt + 1]) + "\n").encode())
            i += 3 + cnt
        elif t == "$Nodes":
            cnt = int(L[i + 1])
            out.append(f"$Nodes\n{cnt}\n".encode())
            for r in range(cnt):
                p = L[i + 2 + r].split()
                out += [pk(i32, [int(p[0])]), pk(f64, [float(v) for v in p[1:4]])]
            out.append(b"\n$EndNodes\n")
            i += 3 + cnt
        elif t == "$Elements":
            cnt = int(L[i + 1])
            out.append(f"$Elements\n{cnt}\n".encode())
            for r in range(cnt):
                p = [int(v) for v in L[i + 2 + r].split()]
                num, et, ntag = p[0], p[1], p[2]
                out += [pk(i32, [et, 1, ntag]), pk(i32, [num] + p[3:])]
            out.append(b"\n$EndElements\n")
            i += 3 + cnt
        else:
            i += 1
    return b"".join(out)


def _same_mesh(a, b):
    assert a.order == b.order
    assert np.array_equal(a.lines, b.lines)
    assert list(a.physical) == list(b.physical)
    assert np.array_equal(a.region, b.region)
    assert len(a.cells) == len(b.cells)
    for x, y in zip(a.cells, b.cells):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("endian", ["<", ">"])
def test_binary_41_equals_ascii_fixture(endian, tmp_path):
    src = os.path.join(GOLDEN, "square_2x2_v41.msh")
    ref = B.read_msh(src, name="m")
    path = tmp_path / "m.msh"
    path.write_bytes(_encode_41_binary(open(src).read(), endian))
    _same_mesh(B.read_msh(str(path), name="m"), ref)


@pytest.mark.parametrize("endian", ["<", ">"])
@pytest.mark.parametrize("builder", ["square_p2", "star_p2"])
def test_binary_22_equals_ascii(builder, endian, tmp_path):
    mesh = {"square_p2": lambda: B.quadrado(ndiv=5, ordem=2),
            "star_p2": lambda: B.star_mesh(40, seed=4)}[builder]()
    a = tmp_path / "a.msh"
    B.write_msh(mesh, str(a), fmt="2.2")
    ref = B.read_msh(str(a), name="m")
    b = tmp_path / "b.msh"
    b.write_bytes(_encode_22_binary(a.read_text(), endian))
    _same_mesh(B.read_msh(str(b), name="m"), ref)


@pytest.mark.parametrize("fmt,binary", [("2.2", False), ("4.1", False), ("4.1", True)])
@pytest.mark.parametrize("builder", ["square_p1", "square_p2", "star_p2"])
def test_write_read_all_formats(builder, fmt, binary, tmp_path):
    mesh = {"square_p1": lambda: B.quadrado(ndiv=6),
            "square_p2": lambda: B.quadrado(ndiv=5, ordem=2),
            "star_p2": lambda: B.star_mesh(40, seed=4)}[builder]()
    path = str(tmp_path / "m.msh")
    B.write_msh(mesh, path, fmt=fmt, binary=binary)
    back = B.read_msh(path)
    d0 = B.format2d(mesh, B.Laplace(1.0))
    d1 = B.format2d(back, B.Laplace(1.0))
    assert (d0.ne, d0.n, d0.ni) == (d1.ne, d1.n, d1.ni)
    assert np.allclose(d0.collocation, d1.collocation, rtol=0, atol=1e-13)
    assert np.array_equal(d0.BC, d1.BC) and np.allclose(d0.BV, d1.BV)
    if binary:   # binary coordinates are the doubles themselves
        assert np.array_equal(d0.collocation, d1.collocation)


def test_unknown_sections_are_skipped(tmp_path):
    src = open(os.path.join(GOLDEN, "square_2x2_v41.msh")).read()
    extra = "$Comments\nanything $Nodes here\n$EndComments\n"
    p = tmp_path / "m.msh"
    p.write_bytes(_encode_41_binary(src, "<").replace(b"$PhysicalNames",
                                                      extra.encode() + b"$PhysicalNames", 1))
    _same_mesh(B.read_msh(str(p), name="m"), B.read_msh(os.path.join(GOLDEN, "square_2x2_v41.msh"),
                                                         name="m"))


def test_bad_format_errors(tmp_path):
    p = tmp_path / "m.msh"
    p.write_bytes(b"$MeshFormat\n4.1 1 4\n\x01\x00\x00\x00\n$EndMeshFormat\n")
    with pytest.raises(ValueError):
        B.read_msh(str(p))
    p.write_bytes(b"$Nodes\n0\n$EndNodes\n")
    with pytest.raises(ValueError):
        B.read_msh(str(p))
    with pytest.raises(ValueError):
        B.write_msh(B.quadrado(ndiv=3), str(p), fmt="3.0")
