"""The reference runner's file formats (oracle/refrun.py), without running
the reference: 'ori' checkpoints round-trip and carry the reference's record
layout, the record conversion is the inverse of the oracle's checkpoint
loader, and the namelist holds the groups and settings the reference reads."""
import os

import numpy as np

from oracle import refrun
from oracle.pyoracle import Oracle


def test_ori_round_trip_and_record_markers(tmp_path):
    rng = np.random.default_rng(3)
    rec = rng.standard_normal((3, 4, 5, 39))
    path = os.path.join(tmp_path, "c.bin")
    refrun.write_ori(path, rec)
    raw = np.fromfile(path, dtype="<i4")
    # one Fortran unformatted record per site: length, 39 doubles, length
    per = 2 + 2 * 39
    assert raw.size == 3 * 4 * 5 * per
    assert np.all(raw.reshape(-1, per)[:, 0] == 8 * 39) and np.all(raw.reshape(-1, per)[:, -1] == 8 * 39)
    # x outermost, z innermost
    first = np.fromfile(path, dtype="<f8", count=1 + 39)[0:]
    assert np.array_equal(np.frombuffer(first.tobytes()[4:4 + 8 * 39], dtype="<f8"), rec[0, 0, 0])
    np.testing.assert_array_equal(refrun.read_ori(path, (3, 4, 5), 39), rec)


def test_records_invert_the_oracle_checkpoint_loader():
    for ncomp, r in ((2, 39), (3, 64)):
        o = Oracle(5, 4, 3, ncomp=ncomp)
        o.init_noise(11, amp=(0.01, 0.01, 0.01), dipoles=ncomp == 3)
        rec = refrun.records_from_sites(o.interior(), ncomp == 3)
        assert rec.shape == (5, 4, 3, r)
        o2 = Oracle(5, 4, 3, ncomp=ncomp)
        o2.load_checkpoint_array(rec)
        np.testing.assert_array_equal(o2.populations(), o.populations())
        np.testing.assert_array_equal(refrun.records_from_sites(o2.interior(), ncomp == 3), rec)


def test_namelist_groups_and_settings():
    t = refrun.namelist("x", (8, 9, 10), 7, True, restore=True, n_checkpoint=7, lees_edwards=True, shear_u=0.05)
    for group in ("&fixed_input", "&variable_input", "&lbe_input"):
        assert group in t
    for line in ("nx = 8", "ny = 9", "nz = 10", "n_iteration = 7", "init_cond = 7", "n_checkpoint = 7",
                 "checkpoint_format = 'ori'", "boundary_cond = -1", "inv_fluid = 5", "shear_u = 0.05",
                 "restore_string = 't00000000-0000000000'"):
        assert line in t, line
    b = refrun.namelist("y", (4, 4, 4), 3, False)
    assert "init_cond = -1" in b and "g_bs = 0.0" in b and "fg = 0.0" in b
