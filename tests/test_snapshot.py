"""Snapshot I/O (restart files) on host vectors (no GPU): bit-exact round trips in FP64 and
FP32 with the discretization parameters, rotation of older snapshots as the reference keeps
them (restart.h:57-72), and refusal of corrupted, truncated or mismatched files before any
destination is written."""
import os

import numpy as np
import pytest

import exadg_amd


def _vectors(n, k, dtype, seed=0):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(n).astype(dtype) for _ in range(k)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_round_trip_is_bit_exact(tmp_path, dtype):
    path = str(tmp_path / "seq.snapshot")
    vs = _vectors(1001, 3, dtype)
    mesh = exadg_amd.mesh_desc((4, 5, 6), deformation=0.1, mapping_degree=5, periodic=(1, 1, 0))
    exadg_amd.snapshot_write(path, vs, time=0.37, step=37, degree=5, mesh=mesh,
                             user=[1.5, -2.0, 3e-7])
    info = exadg_amd.snapshot_read_info(path)
    assert info["n"] == 1001 and info["n_vectors"] == 3 and info["degree"] == 5
    assert info["time"] == 0.37 and info["step"] == 37 and info["user"][:3] == [1.5, -2.0, 3e-7]
    assert info["mesh"]["n_cells"] == [4, 5, 6] and info["mesh"]["periodic"] == [1, 1, 0]
    assert info["mesh"]["deformation"] == 0.1 and info["mesh"]["mapping_degree"] == 5
    out = [np.full(1001, np.nan, dtype=dtype) for _ in range(3)]
    exadg_amd.snapshot_read(path, out)
    for a, b in zip(vs, out):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_old_snapshots_are_rotated(tmp_path):
    path = str(tmp_path / "run.snapshot")
    for step in range(4):
        exadg_amd.snapshot_write(path, _vectors(10, 1, np.float64, seed=step), step=step, keep=3)
    assert exadg_amd.snapshot_read_info(path)["step"] == 3
    assert exadg_amd.snapshot_read_info(path + ".old.1")["step"] == 2
    assert exadg_amd.snapshot_read_info(path + ".old.2")["step"] == 1
    assert not os.path.exists(path + ".old.3") and not os.path.exists(path + ".tmp")


def test_corrupted_truncated_and_mismatched_files_are_refused(tmp_path):
    path = str(tmp_path / "a.snapshot")
    vs = _vectors(500, 2, np.float64)
    exadg_amd.snapshot_write(path, vs)
    raw = bytearray(open(path, "rb").read())
    out = [np.zeros(500) for _ in range(2)]
    # one flipped bit in the data
    bad = str(tmp_path / "bad.snapshot")
    raw2 = bytearray(raw)
    raw2[len(raw2) // 2] ^= 0x10
    open(bad, "wb").write(raw2)
    with pytest.raises(exadg_amd.ExaDGError, match="checksum"):
        exadg_amd.snapshot_read(bad, out)
    assert all(not v.any() for v in out)  # nothing written from the corrupted file
    # truncated
    open(bad, "wb").write(raw[:-20])
    with pytest.raises(exadg_amd.ExaDGError, match="truncated"):
        exadg_amd.snapshot_read(bad, out)
    # not a snapshot
    open(bad, "wb").write(b"hello world, not a snapshot")
    with pytest.raises(exadg_amd.ExaDGError, match="not a snapshot"):
        exadg_amd.snapshot_read_info(bad)
    # expectations that differ from the file
    with pytest.raises(exadg_amd.ExaDGError, match="holds 2 vectors"):
        exadg_amd.snapshot_read(path, [np.zeros(500)])
    with pytest.raises(exadg_amd.ExaDGError, match="holds 2 vectors"):
        exadg_amd.snapshot_read(path, [np.zeros(499), np.zeros(499)])
    with pytest.raises(exadg_amd.ExaDGError, match="precision"):
        exadg_amd.snapshot_read(path, [np.zeros(500, np.float32), np.zeros(500, np.float32)])
    with pytest.raises(exadg_amd.ExaDGError, match="can not open"):
        exadg_amd.snapshot_read_info(str(tmp_path / "missing.snapshot"))
