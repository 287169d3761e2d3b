"""The reference's file formats (host-only library cbns3d_cuda_amd/libcbns_io.so):
legacy VTK text (src/Block3d.cpp:357-409), Plot3D ASCII (:151-202), binary
checkpoint (:411-474) and the .ini / positional input.txt parameters
(:44-92, python/rae2822.ini, python/parser_ini.py)."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
IOLIB = os.path.join(ROOT, "cbns3d_cuda_amd", "libcbns_io.so")
D = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def io():
    if not os.path.exists(IOLIB):
        import subprocess
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "cbns3d_cuda_amd"), "libcbns_io.so"],
                       check=True)
    L = C.CDLL(IOLIB)
    L.cbns_io_write_vtk.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int] + [D] * 8
    L.cbns_io_write_plot3d.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, D, D, D]
    L.cbns_io_read_plot3d_dims.argtypes = [C.c_char_p, C.POINTER(C.c_int)]
    L.cbns_io_read_plot3d.argtypes = [C.c_char_p, D, D, D]
    L.cbns_io_write_checkpoint.argtypes = [C.c_char_p, C.c_double, C.c_uint32, C.c_uint32,
                                           C.c_uint32, D]
    L.cbns_io_read_checkpoint.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32, D, D]
    L.cbns_io_read_input.argtypes = [C.c_char_p, D, C.POINTER(C.c_int), C.c_char_p, C.c_char_p,
                                     C.c_int]
    return L


def p(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(D)


def test_vtk_text_is_the_reference_format(io, tmp_path):
    IM, JM, KM = 3, 2, 2
    rng = np.random.default_rng(0)
    x, y, z = (rng.standard_normal(IM * JM * KM) for _ in range(3))
    g = (KM + 4, JM + 4, IM + 4)
    fields = [rng.standard_normal(g) for _ in range(5)]
    path = str(tmp_path / "out.vtk")
    assert io.cbns_io_write_vtk(path.encode(), IM, JM, KM, p(x), p(y), p(z), *(p(f) for f in fields)) == 0
    n = IM * JM * KM
    exp = ["# vtk DataFile Version 3.0", "vtk output", "ASCII", "DATASET STRUCTURED_GRID",
           "DIMENSIONS %d %d %d" % (IM, JM, KM), "POINTS %d double" % n]
    exp += ["%E %E %E" % t for t in zip(x, y, z)]
    exp += ["POINT_DATA %d" % n, "FIELD FieldData 5"]
    for name, f in zip(("rho", "u", "v", "w", "p"), fields):
        exp.append("%s 1 %d double" % (name, n))
        exp += ["%E" % f[k + 2, j + 2, i + 2] for k in range(KM) for j in range(JM) for i in range(IM)]
    assert open(path).read() == "\n".join(exp) + "\n"


def test_plot3d_round_trip_and_reference_layout(io, tmp_path):
    IM, JM, KM = 4, 3, 2
    rng = np.random.default_rng(1)
    x, y, z = (rng.standard_normal(IM * JM * KM) for _ in range(3))
    path = str(tmp_path / "grid.xyz")
    assert io.cbns_io_write_plot3d(path.encode(), IM, JM, KM, p(x), p(y), p(z)) == 0
    toks = open(path).read().split()
    assert [int(t) for t in toks[:3]] == [IM, JM, KM]
    assert np.array_equal(np.array(toks[3:], dtype=float), np.concatenate([x, y, z]))
    # any whitespace layout reads (the reference reads with operator>>)
    path2 = str(tmp_path / "grid2.xyz")
    with open(path2, "w") as f:
        f.write("%d %d %d\n" % (IM, JM, KM))
        for a in (x, y, z):
            f.write(" ".join("%.17g" % v for v in a) + "\n")
    dims = (C.c_int * 3)()
    assert io.cbns_io_read_plot3d_dims(path2.encode(), dims) == 0 and list(dims) == [IM, JM, KM]
    ox, oy, oz = (np.zeros(IM * JM * KM) for _ in range(3))
    assert io.cbns_io_read_plot3d(path2.encode(), p(ox), p(oy), p(oz)) == 0
    assert np.array_equal(ox, x) and np.array_equal(oy, y) and np.array_equal(oz, z)
    with open(path2, "w") as f:
        f.write("%d %d %d\n1 2 3\n" % (IM, JM, KM))
    assert io.cbns_io_read_plot3d(path2.encode(), p(ox), p(oy), p(oz)) != 0  # short file


def test_checkpoint_bytes_and_errors(io, tmp_path):
    IMG, JMG, KMG = 7, 6, 5
    Q = np.random.default_rng(2).standard_normal(5 * IMG * JMG * KMG)
    path = str(tmp_path / "checkpoint.dat")
    assert io.cbns_io_write_checkpoint(path.encode(), 1.25, IMG, JMG, KMG, p(Q)) == 0
    raw = open(path, "rb").read()
    assert len(raw) == 8 + 12 + 8 * Q.size
    t, a, b, c = struct.unpack("=d3I", raw[:20])
    assert (t, a, b, c) == (1.25, IMG, JMG, KMG)
    assert np.array_equal(np.frombuffer(raw[20:], dtype=np.float64), Q)
    out = np.zeros_like(Q)
    tc = C.c_double(0)
    assert io.cbns_io_read_checkpoint(path.encode(), IMG, JMG, KMG, C.byref(tc), p(out)) == 0
    assert tc.value == 1.25 and np.array_equal(out, Q)
    assert io.cbns_io_read_checkpoint(path.encode(), IMG + 1, JMG, KMG, C.byref(tc), p(out)) != 0
    open(path, "wb").write(raw[:100])
    assert io.cbns_io_read_checkpoint(path.encode(), IMG, JMG, KMG, C.byref(tc), p(out)) != 0


RAE_INI = """[string]
grid_fname = grid.xyz
checkpoint_fname = checkpoint.dat

[real]
CFL = 0.8
t_cur = 0.0
t_end = 100.0

Pr = 0.71
Pr_t = 0.9
gamma = 1.4

Mach = 0.729
Re = 6.5e6
T_inf = 255.556
angle_attack = 2.31

[integer]
nstep_max = 1000000
checkpoint_freq = 1000

i_begin = 17
i_end = 117

[bool]
is_restart = 1
"""


def _read(io, path):
    num = (C.c_double * 19)()
    ints = (C.c_int * 10)()
    g = C.create_string_buffer(256)
    c = C.create_string_buffer(256)
    rc = io.cbns_io_read_input(path.encode(), num, ints, g, c, 256)
    return rc, list(num), list(ints), g.value.decode(), c.value.decode()


def test_input_ini_and_positional_agree(io, tmp_path):
    ini = tmp_path / "rae2822.ini"
    ini.write_text(RAE_INI)
    rc, num, ints, g, c = _read(io, str(ini))
    assert rc == 0 and g == "grid.xyz" and c == "checkpoint.dat"
    assert num[:10] == [0.8, 0.0, 100.0, 0.71, 0.9, 1.4, 0.729, 6.5e6, 255.556, 2.31]
    assert ints[:5] == [1000000, 1000, 17, 117, 1]
    assert ints[5:] == [1, 0, 0, 0, 0]  # defaults of the added keys
    # what python/parser_ini.py writes: the values, one per line, in file order
    vals = [ln.split("=", 1)[1].strip() for ln in RAE_INI.splitlines() if "=" in ln]
    txt = tmp_path / "input.txt"
    txt.write_text("\n".join(vals) + "\n")
    assert _read(io, str(txt)) == (rc, num, ints, g, c)
    txt.write_text("\n".join(vals[:10]) + "\n")
    assert _read(io, str(txt))[0] != 0  # short input.txt


def test_input_added_keys(io, tmp_path):
    ini = tmp_path / "tgv.ini"
    ini.write_text(RAE_INI + "\n[case]\nviscous = 1\nvisc_centered = 1\nbc_xi = periodic\n"
                   "bc_eta = periodic\nbc_zeta = periodic\nperiod_xi = 6.5 0 0\n"
                   "period_eta = 0 6.5 0\nperiod_zeta = 0 0 6.5\n")
    rc, num, ints, g, c = _read(io, str(ini))
    assert rc == 0 and ints[5:] == [1, 1, 1, 1, 1]
    assert num[10:] == [6.5, 0, 0, 0, 6.5, 0, 0, 0, 6.5]
