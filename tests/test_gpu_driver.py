"""The cbns3d driver (Block3d::solve flow, src/Block3d.cpp:204-313) on the
reference's own file formats: Plot3D mesh + .ini in, checkpoint + VTK out,
freestream initial condition and the reference's boundary rules.  The final
checkpoint is compared with the oracle running the same steps; a restart
from a checkpoint continues bit for bit."""
import os
import struct
import subprocess

import numpy as np
import pytest

from cbns3d_cuda_amd import cases
from helpers import G, oracle_params

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cbns3d_cuda_amd", "cbns3d")


def write_plot3d(path, x, y, z):
    KM, JM, IM = x.shape
    with open(path, "w") as f:
        f.write("%d %d %d\n" % (IM, JM, KM))
        for a in (x, y, z):
            f.write("\n".join("%.17g" % v for v in a.ravel()) + "\n")


def write_ini(path, case, nstep, restart, t_end=100.0, ckpt="checkpoint.dat"):
    p = case.params
    with open(path, "w") as f:
        f.write("[string]\ngrid_fname = grid.xyz\ncheckpoint_fname = %s\n" % ckpt)
        f.write("[real]\nCFL = %r\nt_cur = 0.0\nt_end = %r\nPr = %r\nPr_t = 0.9\ngamma = %r\n"
                % (p.CFL, t_end, p.Pr, p.gamma))
        f.write("Mach = %r\nRe = %r\nT_inf = %r\nangle_attack = %r\n" % (p.Mach, p.Re, p.T_inf,
                                                                         p.angle_attack))
        f.write("[integer]\nnstep_max = %d\ncheckpoint_freq = 1000\ni_begin = %d\ni_end = %d\n"
                % (nstep, p.i_begin, p.i_end))
        f.write("[bool]\nis_restart = %d\n[case]\nviscous = %d\n" % (1 if restart else 0, p.viscous))


def read_ckpt(path, shape):
    raw = open(path, "rb").read()
    t, a, b, c = struct.unpack("=d3I", raw[:20])
    assert (c, b, a) == shape
    return t, np.frombuffer(raw[20:], dtype=np.float64).reshape(shape + (5,))


def run(workdir, ini):
    print("driver", ini, flush=True)
    r = subprocess.run([DRIVER, ini], cwd=workdir, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def oracle_run(po, case, nsteps):
    o = po.Oracle(oracle_params(po, case))
    o.set_mesh(case.x, case.y, case.z, None)
    g = o.g
    aoa = g.angle_attack * 3.14159265358979323846 / 180.0
    o.prim[0], o.prim[1], o.prim[2], o.prim[3] = 1.0, np.cos(aoa), np.sin(aoa), 0.0
    o.prim[4], o.prim[5] = g.p_inf, 1.0
    o.conservative()
    o.Qp[...] = o.Q
    o.bc_primitive()
    t = 0.0
    for _ in range(nsteps):
        t, _ = o.rk3_step(100.0, t)
    return t, o


@pytest.mark.parametrize("viscous", [0, 1])
def test_driver_matches_oracle(tmp_path, oracle, viscous):
    case = cases.c_grid_channel(viscous=viscous)
    write_plot3d(tmp_path / "grid.xyz", case.x, case.y, case.z)
    write_ini(tmp_path / "run.ini", case, nstep=3, restart=True)
    out = run(str(tmp_path), "run.ini")
    assert "The output is successful!" in out and "Elapsed Time = " in out
    assert len([l for l in out.splitlines() if l.count(",") == 2 and l[0].isdigit()]) >= 3
    p = case.params
    shape = (p.KM + 4, p.JM + 4, p.IM + 4)
    t, q = read_ckpt(tmp_path / "checkpoint.dat", shape)
    t_o, o = oracle_run(oracle, case, 3)
    assert t == pytest.approx(t_o, rel=1e-12)
    ref = np.moveaxis(o.Q[:, G - 2:G + p.KM + 2, G - 2:G + p.JM + 2, G - 2:G + p.IM + 2], 0, -1)
    err = np.abs(q - ref).max(axis=(0, 1, 2)) / np.abs(ref).max(axis=(0, 1, 2))
    assert err.max() < 1e-10, err
    vtk = open(tmp_path / "output.vtk").read().splitlines()
    assert vtk[4] == "DIMENSIONS %d %d %d" % (p.IM, p.JM, p.KM)


def test_driver_restart_continues_bit_for_bit(tmp_path):
    case = cases.c_grid_channel(viscous=0)
    write_plot3d(tmp_path / "grid.xyz", case.x, case.y, case.z)
    p = case.params
    shape = (p.KM + 4, p.JM + 4, p.IM + 4)
    write_ini(tmp_path / "a.ini", case, nstep=3, restart=True, ckpt="a.dat")
    run(str(tmp_path), "a.ini")
    write_ini(tmp_path / "b.ini", case, nstep=2, restart=True, ckpt="b.dat")
    run(str(tmp_path), "b.ini")
    write_ini(tmp_path / "c.ini", case, nstep=1, restart=False, ckpt="b.dat")
    out = run(str(tmp_path), "c.ini")
    ta, qa = read_ckpt(tmp_path / "a.dat", shape)
    tb, qb = read_ckpt(tmp_path / "b.dat", shape)
    assert ta == tb and np.array_equal(qa, qb), out


def test_driver_stops_at_t_end(tmp_path):
    case = cases.c_grid_channel(viscous=0)
    write_plot3d(tmp_path / "grid.xyz", case.x, case.y, case.z)
    write_ini(tmp_path / "run.ini", case, nstep=50, restart=True, t_end=0.004)
    out = run(str(tmp_path), "run.ini")
    assert "Finished!" in out
    t, _ = read_ckpt(tmp_path / "checkpoint.dat", (case.params.KM + 4, case.params.JM + 4,
                                                   case.params.IM + 4))
    assert t == pytest.approx(0.004, rel=1e-12)
