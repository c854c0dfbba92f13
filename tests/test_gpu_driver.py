"""The drop-in driver (cbns3d_amd) on an unchanged input.txt: its files are
the reference's, byte for byte.  GPU only."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cbns3d_cart_mpi_amd", "bin", "cbns3d_amd")
NG = 4


def input_txt(N, nsteps, num_output=1, freq=None, restart=False, t_end="20.0"):
    vals = ["checkpoint", "0.5", "0.0", t_end, "0.71", "1.4", "0.1", "1600.0", "255.556", "1.0",
            str(N[0]), str(N[1]), str(N[2]), str(nsteps), str(freq or nsteps), str(num_output),
            "0" if restart else "1"]
    return "\n".join(vals) + "\n"


def run_driver(tmp, N, nsteps, dims, extra=(), **kw):
    (tmp / "input.txt").write_text(input_txt(N, nsteps, **kw))
    r = subprocess.run([EXE, *map(str, dims), *extra], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.mark.parametrize("extra", [(), ("--fused",), ("--async-io",), ("--fused", "--async-io")])
def test_driver_files_byte_identical_to_reference(golden, tmp_path, extra):
    g = golden("ref_vtk8.npz")
    out = run_driver(tmp_path, (8, 8, 8), 2, (1, 1, 1), extra)
    for nm in ("output000/output000.vtr", "output000/output.vtm", "output001/output000.vtr",
               "output001/output.vtm", "checkpoint000.dat"):
        key = nm.replace("/", "__").replace(".", "_")
        assert (tmp_path / nm).read_bytes() == bytes(g[key]), nm
    ref_lines = [ln for ln in bytes(g["stdout"]).decode().splitlines() if "Wall" not in ln and "CPU" not in ln]
    got_lines = [ln for ln in out.splitlines() if "Wall" not in ln and "CPU" not in ln]
    assert got_lines == ref_lines


def read_ckpt(fn):
    b = open(fn, "rb").read()
    t = float(np.frombuffer(b[:8], np.float64)[0])
    IG, JG, KG = (int(x) for x in np.frombuffer(b[8:20], np.uint32))
    return t, np.frombuffer(b[20:], np.float64).reshape(KG, JG, IG, 5)


def test_driver_two_blocks_checkpoints_match_reference(golden, tmp_path):
    g = golden("ref_tgv16.npz")
    run_driver(tmp_path, (16, 16, 16), 10, (2, 1, 1))
    t0, Q0 = read_ckpt(tmp_path / "checkpoint000.dat")
    t1, Q1 = read_ckpt(tmp_path / "checkpoint001.dat")
    assert t0 == t1 == float(g["t_n10"])
    full = np.concatenate([Q0[NG:-NG, NG:-NG, NG:-NG], Q1[NG:-NG, NG:-NG, NG:-NG]], axis=2)
    assert np.array_equal(full, g["Q_n10"])


def test_driver_restart_equals_oracle_restart(tmp_path):
    """is_restart = 0 resumes from the driver's own checkpoint with mu = 0 at
    the first dt, exactly as the reference's restart path does."""
    from oracle import oracle as O

    run_driver(tmp_path, (16, 16, 16), 1, (1, 1, 1))
    t1, Q1 = read_ckpt(tmp_path / "checkpoint000.dat")
    vals = input_txt((16, 16, 16), 1).splitlines()
    vals[-1] = "0"
    (tmp_path / "input.txt").write_text("\n".join(vals) + "\n")
    r = subprocess.run([EXE, "1", "1", "1"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    t2, Q2 = read_ckpt(tmp_path / "checkpoint000.dat")
    b = O.Block(O.make_params((16, 16, 16)))
    b.set_state(Q1, t1, mu0=0.0)
    b.run(1)
    assert t2 == b.t_cur
    assert np.array_equal(Q2[NG:-NG, NG:-NG, NG:-NG], b.Q_interior(0))


def test_driver_two_block_restart_uses_the_rank_suffixed_names(tmp_path):
    """The driver forms the checkpoint name as the reference's main does
    (rank suffix, src/cbns3d.cpp:132-134) and the facade maps it to one file
    per block: checkpoint000.dat and checkpoint001.dat, which a restart with
    the same dims reads back (equal to the oracle's restart)."""
    from oracle import oracle as O

    run_driver(tmp_path, (16, 16, 16), 1, (2, 1, 1))
    names = sorted(p.name for p in tmp_path.iterdir() if p.name.startswith("checkpoint"))
    assert names == ["checkpoint000.dat", "checkpoint001.dat"], names
    t1, Q1a = read_ckpt(tmp_path / "checkpoint000.dat")
    _, Q1b = read_ckpt(tmp_path / "checkpoint001.dat")
    Q1 = np.zeros((24, 24, 24, 5))
    Q1[NG:-NG, NG:-NG, NG:-NG] = np.concatenate([Q1a[NG:-NG, NG:-NG, NG:-NG], Q1b[NG:-NG, NG:-NG, NG:-NG]], axis=2)
    vals = input_txt((16, 16, 16), 1).splitlines()
    vals[-1] = "0"
    (tmp_path / "input.txt").write_text("\n".join(vals) + "\n")
    r = subprocess.run([EXE, "2", "1", "1"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    names = sorted(p.name for p in tmp_path.iterdir() if p.name.startswith("checkpoint"))
    assert names == ["checkpoint000.dat", "checkpoint001.dat"], names
    t2, Q2a = read_ckpt(tmp_path / "checkpoint000.dat")
    _, Q2b = read_ckpt(tmp_path / "checkpoint001.dat")
    b = O.Block(O.make_params((16, 16, 16)))
    b.set_state(Q1, t1, mu0=0.0)
    b.run(1)
    assert t2 == b.t_cur
    got = np.concatenate([Q2a[NG:-NG, NG:-NG, NG:-NG], Q2b[NG:-NG, NG:-NG, NG:-NG]], axis=2)
    assert np.array_equal(got, b.Q_interior(0))


def test_driver_mu_sidecar_restart_continues_bit_for_bit(tmp_path):
    """Two blocks: a run of 3 steps writes its checkpoints with mu sidecars;
    a restart from them for 3 more steps ends with the files of an
    uninterrupted 6-step run, byte for byte."""
    a, c = tmp_path / "a", tmp_path / "c"
    a.mkdir()
    c.mkdir()
    run_driver(a, (16, 16, 16), 3, (2, 1, 1), ("--mu-sidecar",))
    assert (a / "checkpoint001.dat.mu").exists()
    run_driver(a, (16, 16, 16), 3, (2, 1, 1), ("--mu-sidecar",), restart=True)
    run_driver(c, (16, 16, 16), 6, (2, 1, 1))
    for nm in ("checkpoint000.dat", "checkpoint001.dat"):
        assert (a / nm).read_bytes() == (c / nm).read_bytes(), nm


def test_driver_async_io_writes_the_files_of_the_synchronous_writers(tmp_path):
    """Checkpoints every 2 steps (one file name per block, rewritten) and
    VTK output of two blocks at every 0.0125 of time (t reaches about 0.078
    in 12 steps: six outputs besides the first and the last): the
    asynchronous writers leave every file as the synchronous ones do."""
    runs = {}
    for name, extra in (("sync", ()), ("async", ("--async-io",))):
        d = tmp_path / name
        d.mkdir()
        run_driver(d, (16, 16, 16), 12, (2, 1, 1), extra, num_output=20, freq=2, t_end="0.25")
        runs[name] = {str(p.relative_to(d)): p.read_bytes() for p in sorted(d.rglob("*")) if p.is_file()
                      and p.name != "input.txt"}
    assert sum(nm.endswith(".vtr") for nm in runs["sync"]) >= 12
    assert runs["sync"].keys() == runs["async"].keys()
    for nm in runs["sync"]:
        assert runs["sync"][nm] == runs["async"][nm], nm
