"""Configuration text format (conf.cc:55-96,143-204, su3.cc:225-243) against
a file the reference's own Conf::write_last wrote and its Conf(init=2) read."""
import gzip
import os

import numpy as np

from rhmc_amd.conf_io import format_conf, read_conf, write_conf

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _ref_text():
    with gzip.open(os.path.join(GOLD, "config_t4.txt.gz"), "rt") as fh:
        return fh.read()


def test_writer_is_byte_identical_to_reference():
    g = np.load(os.path.join(GOLD, "golden_t4.npz"))
    txt = format_conf(g["links"], tuple(g["dims"]), float(g["beta"]), float(g["mass"]), float(g["n_flavours"]), 7)
    assert txt == _ref_text()


def test_reader_matches_reference_reader(tmp_path):
    g = np.load(os.path.join(GOLD, "golden_t4.npz"))
    p = tmp_path / "config"
    p.write_text(_ref_text())
    U, hdr = read_conf(str(p), dims=tuple(g["dims"]))
    assert hdr["update_iteration"] == 7 and hdr["beta"] == float(g["beta"])
    # Conf(init=2) keeps the links as read; both parse the same decimal text
    assert np.array_equal(U, g["conf_read_links"])
    assert np.abs(U - g["links"]).max() < 1e-15


def test_round_trip_and_extent_check(tmp_path):
    g = np.load(os.path.join(GOLD, "golden_t4.npz"))
    p = str(tmp_path / "config_0")
    write_conf(p, g["links"], (4, 4, 4, 4), 5.5, 0.1, 2, 3)
    U, hdr = read_conf(p)
    assert hdr["dims"] == (4, 4, 4, 4) and hdr["update_iteration"] == 3
    assert np.abs(U - g["links"]).max() < 1e-15
    try:
        read_conf(p, dims=(4, 4, 4, 8))
        raise AssertionError("extent mismatch not detected")
    except ValueError:
        pass
