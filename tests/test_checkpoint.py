"""Checkpoints as the reference's restart files (fvg_checkpoint_*; SimulationRestart.cpp:61-85, 162-190, 254-360).

The container is pinned by numpy's own .npz reader and writer (an independent implementation of the format); the
keys, dtypes and rules by the reference's own save_checkpoint / load_checkpoint (oracle/_ref); the hand-over by
runs that continue from a checkpoint bit for bit, on the device and across the reference and the device."""
import os
import zipfile

import numpy as np
import pytest

from cases import VORTEX_HANCOCK_256, base_keys, own_case_file, smooth_state, vortex_state
from fenice_amd import fvg
from fenice_amd.workloads import taylor_green, taylor_green_keys

VORTEX_64 = dict(VORTEX_HANCOCK_256, nx="64", ny="64", cfl="0.5")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def keys_3d():
    return base_keys(6, 5, 4, reconstruction="weno5", time_scheme="SSP Runge-Kutta 3")


def test_checkpoint_is_a_numpy_npz(tmp_path):
    cfg = fvg.config_from_keys(keys_3d())
    U = smooth_state(6, 5, 4, 3)
    path = str(tmp_path / "c.npz")
    fvg.checkpoint_save(cfg, U, 0.125, 17, path, lagged_speed=3.5)
    with np.load(path, allow_pickle=False) as d:
        assert set(d.files) == {"U", "t", "step", "S", "format_version", "nx", "ny", "nz", "Lx", "Ly", "Lz", "collection_time",
                                "collection_file"}
        assert d["U"].dtype == np.dtype("<f8") and d["U"].shape == (120, 5) and np.array_equal(bits(d["U"]), bits(U))
        assert d["t"].shape == () and d["t"].dtype == np.float64 and d["t"] == 0.125
        assert d["step"].dtype == np.int64 and d["step"] == 17 and d["format_version"] == 1 and d["S"] == 3.5
        assert (d["nx"], d["ny"], d["nz"]) == (6, 5, 4) and d["Lx"] == cfg.Lx and d["Lz"] == cfg.Lz
        assert d["collection_time"].shape == (0,) and d["collection_file"].dtype == np.dtype("<U1")
    with zipfile.ZipFile(path) as z:   # stored entries, in key order, each with a good CRC
        assert [i.filename for i in z.infolist()] == sorted(i.filename for i in z.infolist())
        assert all(i.compress_type == zipfile.ZIP_STORED for i in z.infolist()) and z.testzip() is None
    U2, t, step, s = fvg.checkpoint_load(cfg, path)
    assert np.array_equal(bits(U2), bits(U)) and (t, step, s) == (0.125, 17, 3.5)


def test_2d_checkpoint_has_four_columns_and_no_s(tmp_path):
    cfg = fvg.config_from_keys(base_keys(7, 6))
    U = smooth_state(7, 6, 1, 5)[:, :4]
    path = str(tmp_path / "c.npz")
    fvg.checkpoint_save(cfg, U, 1.0, 3, path)
    with np.load(path, allow_pickle=False) as d:
        assert d["U"].shape == (42, 4) and "S" not in d.files
    assert fvg.checkpoint_load(cfg, path)[3] == 0.0


def test_numpy_written_npz_is_read(tmp_path):
    # numpy's writer (zip64 local headers) with a 4-column state on a 3-D mesh: rho w comes back as 0
    cfg = fvg.config_from_keys(keys_3d())
    U = smooth_state(6, 5, 4, 8)[:, :4]
    path = str(tmp_path / "n.npz")
    np.savez(path, U=U, t=np.float64(2.5), step=np.int64(40), S=np.float64(1.25), nx=np.int64(6), format_version=np.int64(1))
    U2, t, step, s = fvg.checkpoint_load(cfg, path)
    assert np.array_equal(bits(U2[:, :4]), bits(U)) and np.all(U2[:, 4] == 0.0) and (t, step, s) == (2.5, 40, 1.25)


def test_zip64_fields_are_written_and_read(tmp_path, monkeypatch):
    cfg = fvg.config_from_keys(keys_3d())
    U = smooth_state(6, 5, 4, 9)
    path = str(tmp_path / "z.npz")
    monkeypatch.setenv("FVG_NPZ_ZIP64_FROM", "64")   # every entry larger than 64 bytes, and offsets past it
    fvg.checkpoint_save(cfg, U, 0.5, 2, path)
    with zipfile.ZipFile(path) as z:
        assert any(i.file_size >= 64 and i.extra for i in z.infolist()) and z.testzip() is None
    with np.load(path, allow_pickle=False) as d:
        assert np.array_equal(bits(d["U"]), bits(U)) and d["step"] == 2
    assert np.array_equal(bits(fvg.checkpoint_load(cfg, path)[0]), bits(U))


def test_bad_checkpoints_are_refused(tmp_path):
    cfg = fvg.config_from_keys(keys_3d())
    U = smooth_state(6, 5, 4, 3)
    good = str(tmp_path / "g.npz")
    fvg.checkpoint_save(cfg, U, 0.0, 0, good)
    other = fvg.config_from_keys(base_keys(6, 5, 5))
    with pytest.raises(fvg.InvalidArgument):   # another mesh
        fvg.checkpoint_load(other, good)
    newer = str(tmp_path / "v2.npz")
    np.savez(newer, U=U, t=np.float64(0), step=np.int64(0), format_version=np.int64(2))
    with pytest.raises(fvg.CheckpointIOError):
        fvg.checkpoint_load(cfg, newer)
    packed = str(tmp_path / "p.npz")
    np.savez_compressed(packed, U=U, t=np.float64(0), step=np.int64(0))
    with pytest.raises(fvg.CheckpointIOError):
        fvg.checkpoint_load(cfg, packed)
    raw = open(good, "rb").read()
    cut = str(tmp_path / "cut.npz")
    open(cut, "wb").write(raw[: len(raw) // 2])
    with pytest.raises(fvg.CheckpointIOError):
        fvg.checkpoint_load(cfg, cut)
    flip = bytearray(raw)
    flip[len(raw) // 3] ^= 0x40   # inside U's data: the CRC no longer matches
    bad = str(tmp_path / "crc.npz")
    open(bad, "wb").write(bytes(flip))
    with pytest.raises(fvg.CheckpointIOError):
        fvg.checkpoint_load(cfg, bad)
    with pytest.raises(fvg.CheckpointIOError):
        fvg.checkpoint_load(cfg, str(tmp_path / "missing.npz"))


def test_malformed_central_directory_is_refused(tmp_path, monkeypatch):
    # a central-directory entry whose name, extra field or zip64 field would reach past the directory: refused as an
    # input error, without reading beyond it
    cfg = fvg.config_from_keys(keys_3d())
    U = smooth_state(6, 5, 4, 3)
    good = str(tmp_path / "g.npz")
    fvg.checkpoint_save(cfg, U, 0.0, 0, good)
    raw = open(good, "rb").read()
    last = raw.rfind(b"PK\x01\x02")   # the last entry of the central directory
    long_name = bytearray(raw)
    long_name[last + 28:last + 30] = (0xFFFF).to_bytes(2, "little")
    path = str(tmp_path / "name.npz")
    open(path, "wb").write(bytes(long_name))
    with pytest.raises(fvg.CheckpointIOError):
        fvg.checkpoint_load(cfg, path)
    long_extra = bytearray(raw)
    long_extra[last + 30:last + 32] = (0xFFFF).to_bytes(2, "little")
    path = str(tmp_path / "extra.npz")
    open(path, "wb").write(bytes(long_extra))
    with pytest.raises(fvg.CheckpointIOError):
        fvg.checkpoint_load(cfg, path)
    # zip64 entries (forced on a small file): a zip64 extra field that claims fewer bytes than its sizes need
    monkeypatch.setenv("FVG_NPZ_ZIP64_FROM", "1")
    z64 = str(tmp_path / "z64.npz")
    fvg.checkpoint_save(cfg, U, 0.0, 0, z64)
    raw = bytearray(open(z64, "rb").read())
    c = raw.rfind(b"PK\x01\x02")
    nl = int.from_bytes(raw[c + 28:c + 30], "little")
    x = c + 46 + nl
    assert int.from_bytes(raw[x:x + 2], "little") == 0x0001
    raw[x + 2:x + 4] = (4).to_bytes(2, "little")
    path = str(tmp_path / "short64.npz")
    open(path, "wb").write(bytes(raw))
    with pytest.raises(fvg.CheckpointIOError):
        fvg.checkpoint_load(cfg, path)


def test_reference_reads_our_checkpoint_and_we_read_its(ref, tmp_path):
    # both directions against the reference's own writer and reader, on the MUSCL-Hancock vortex (S is the lagged
    # speed): the reference steps, our writer saves its state, a fresh reference run loads it and continues with the
    # same bits as the original run; the reference's file is read back by our reader exactly
    keys = VORTEX_64
    cfg = fvg.config_from_keys(keys)
    run = ref.create(own_case_file(tmp_path), **keys)
    run.set_state(vortex_state(64))
    t = 0.0
    for _ in range(6):
        t += run.step(t, 1e9)[0]
    U = run.get_state()
    S = run.lagged_speed
    assert S > 0.0
    fvg.checkpoint_save(cfg, U[:, :4], t, 6, run.checkpoint_path, lagged_speed=S)
    fresh = ref.create(own_case_file(tmp_path), **keys)
    t2, step2 = fresh.load_checkpoint()
    assert (t2, step2) == (t, 6) and fresh.lagged_speed == S
    assert np.array_equal(bits(fresh.get_state()), bits(U))
    for _ in range(4):
        dta, _ = run.step(t, 1e9)
        dtb, _ = fresh.step(t2, 1e9)
        assert dta == dtb
        t, t2 = t + dta, t2 + dtb
    assert np.array_equal(bits(fresh.get_state()), bits(run.get_state()))
    run.save_checkpoint(t, 10)   # the reference's writer, our reader
    U3, t3, step3, s3 = fvg.checkpoint_load(cfg, run.checkpoint_path)
    assert np.array_equal(bits(U3), bits(run.get_state()[:, :4])) and (t3, step3, s3) == (t, 10, run.lagged_speed)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["tgv", "vortex"])
def test_device_resume_is_bitwise(tmp_path, case):
    # N steps, checkpoint, a fresh context resumes from the file and takes M more: the same bits as N + M steps in
    # one context (the vortex runs MUSCL-Hancock, whose lagged speed S rides in the file)
    if case == "tgv":
        keys, U0, n1, n2 = taylor_green_keys(24), taylor_green(24), 4, 4
    else:
        keys, U0, n1, n2 = VORTEX_64, vortex_state(64)[:, :4], 6, 6
    cfg = fvg.config_from_keys(keys)
    path = str(tmp_path / "r.npz")
    with fvg.Solver(cfg) as s:
        s.upload(U0)
        t, _, _ = s.advance(0.0, 1e9, max_steps=n1)
        s.checkpoint_write(path, t, n1)
        t, _, _ = s.advance(t, 1e9, max_steps=n2)
        straight = s.download()
    with fvg.Solver(cfg) as s:
        t1, step = s.checkpoint_read(path)
        assert step == n1
        t2, _, _ = s.advance(t1, 1e9, max_steps=n2)
        resumed = s.download()
    assert t2 == t and np.array_equal(bits(resumed), bits(straight))


@pytest.mark.gpu
def test_device_and_reference_hand_over(tmp_path, ref):
    # reference -> device: the reference writes its restart file after 6 MUSCL-Hancock steps, the device reads it
    # and takes 4 more, bit for bit the reference's own next 4; device -> reference: the device writes, a fresh
    # reference run loads the file and its next steps are the device's
    keys = VORTEX_64
    cfg = fvg.config_from_keys(keys)
    run = ref.create(own_case_file(tmp_path), **keys)
    run.set_state(vortex_state(64))
    t = 0.0
    for _ in range(6):
        t += run.step(t, 1e9)[0]
    run.save_checkpoint(t, 6)
    with fvg.Solver(cfg) as s:
        td, step = s.checkpoint_read(run.checkpoint_path)
        assert (td, step) == (t, 6)
        for _ in range(4):
            dd = s.step(td, 1e9).dt
            dr, _ = run.step(t, 1e9)
            assert dd == dr
            td, t = td + dd, t + dr
        assert np.array_equal(bits(s.download()[:, :4]), bits(run.get_state()[:, :4]))
        s.checkpoint_write(str(tmp_path / "dev.npz"), td, 10)
        os.replace(str(tmp_path / "dev.npz"), run.checkpoint_path)
        fresh = ref.create(own_case_file(tmp_path), **keys)
        tf, stepf = fresh.load_checkpoint()
        assert (tf, stepf) == (td, 10) and fresh.lagged_speed == s.lagged_speed
        for _ in range(3):
            dd = s.step(td, 1e9).dt
            dr, _ = fresh.step(tf, 1e9)
            assert dd == dr
            td, tf = td + dd, tf + dr
        assert np.array_equal(bits(s.download()[:, :4]), bits(fresh.get_state()[:, :4]))
