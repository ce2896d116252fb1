"""Binary checkpoints (fvdg_checkpoint_write / fvdg_checkpoint_read): the raw FP64 state with a
header, so a restart continues bit for bit (the reference restarts from text files by projection,
examples/Vortex/tools.h:73-129, which loses digits).  Checked: a bit-exact round trip with time and
step; the header alone; every corruption is a status code (flipped data byte -> checksum, truncated
file, foreign file, wrong sizes); on the GPU, 6 steps equal 3 steps + checkpoint + restart + 3 steps."""
import numpy as np
import pytest

import fve_dg_amd as F
from conftest import VORTEX


@pytest.fixture
def state():
    mesh = F.Mesh.box(3, ub=(10, 10, 10), periodic=7, jitter=0.1, seed=2)
    return mesh, mesh.project(3, "vortex", VORTEX)


def test_round_trip_bit_exact(tmp_path, state):
    _, U = state
    path = tmp_path / "s.ckpt"
    F.checkpoint_write(path, 3, U, t=0.125, step=42)
    V, order, t, step = F.checkpoint_read(path)
    assert order == 3 and t == 0.125 and step == 42
    assert V.dtype == np.float64 and np.array_equal(V, U)
    assert not (tmp_path / "s.ckpt.tmp").exists()
    assert path.stat().st_size == 56 + U.size * 8


def test_corruptions_are_errors(tmp_path, state):
    _, U = state
    path = tmp_path / "s.ckpt"
    F.checkpoint_write(path, 3, U)
    raw = bytearray(path.read_bytes())
    bad = tmp_path / "bad.ckpt"
    flipped = bytearray(raw)
    flipped[56 + 1234] ^= 0x10  # one bit of one coefficient
    bad.write_bytes(bytes(flipped))
    with pytest.raises(F.FvdgError) as e:
        F.checkpoint_read(bad)
    assert e.value.status == F.FVDG_ERR_STATE
    bad.write_bytes(bytes(raw[:-8]))  # truncated
    with pytest.raises(F.FvdgError) as e:
        F.checkpoint_read(bad)
    assert e.value.status == F.FVDG_ERR_STATE
    bad.write_bytes(b"not a checkpoint at all, but long enough for a header........")
    with pytest.raises(F.FvdgError) as e:
        F.checkpoint_read(bad)
    assert e.value.status == F.FVDG_ERR_ARG
    with pytest.raises(ValueError):
        F.checkpoint_write(path, 2, U)  # 5N of order 2 does not match the state
    with pytest.raises(F.FvdgError):
        F.checkpoint_read(tmp_path / "missing.ckpt")


@pytest.mark.gpu
def test_restart_continues_bit_for_bit(tmp_path, solver, state):
    print("[case] checkpoint restart", flush=True)
    mesh, U0 = state
    solver.set_scheme(3, flux="roe").set_mesh(mesh)
    solver.upload(U0)
    dt = solver.stable_dt(0.3)
    solver.ssprk3(dt, 6)
    straight = solver.download()
    solver.upload(U0)
    solver.ssprk3(dt, 3)
    F.checkpoint_write(tmp_path / "mid.ckpt", 3, solver.download(), t=3 * dt, step=3)
    other = F.Solver(0)
    try:
        V, order, t, step = F.checkpoint_read(tmp_path / "mid.ckpt")
        other.set_scheme(order, flux="roe").set_mesh(mesh)
        other.upload(V)
        other.ssprk3(dt, 6 - step)
        assert np.array_equal(other.download(), straight)
    finally:
        other.close()
