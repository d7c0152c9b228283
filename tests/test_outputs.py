"""The output writers (libpic_io.so, include/pic_io.h) against the files the
reference's own Output functions wrote (Outputs.cpp:9-205) for the same
data: the VTK ImageData of the node fields, the PolyData of sampled
particles per species, and the runtime diagnostics CSV, compared as text
(the CSV's wall-clock column aside).  No GPU."""
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def texts():
    return np.load(os.path.join(ROOT, "tests", "golden", "ref_outputs.npz"))


def _text(texts, name):
    return bytes(texts[name]).decode()


def test_vti_matches_reference(golden, texts, tmp_path):
    import pic_io

    gr = golden["out_grid"]
    ni, nj, nk = int(gr[0]), int(gr[1]), int(gr[2])
    species = []
    for s, name in enumerate(("e-", "O+")):
        t = str(s)
        species.append({"name": name, "den": golden["out_den_" + t], "den_avg": golden["out_den_avg_" + t],
                        "vel": golden["out_vel_" + t], "T": golden["out_T_" + t],
                        "mpc": golden["out_mpc_" + t]})
    path = str(tmp_path / "fields_00002.vti")
    pic_io.write_vti(path, ni, nj, nk, gr[3:6], gr[6:9], golden["out_node_vol"],
                     golden["out_object_id"], golden["out_node_type"], golden["out_phi"],
                     golden["out_rho"], golden["out_ef"], species)
    got = open(path).read()
    want = _text(texts, "fields_00002.vti")
    assert "ObjectID" in want and "mpc.O+" in want
    assert got == want


@pytest.mark.parametrize("s,name", [(0, "e-"), (1, "O+")])
def test_vtp_matches_reference(golden, texts, tmp_path, s, name):
    import pic_io

    path = str(tmp_path / ("parts_%s.vtp" % name))
    pic_io.write_vtp(path, name, golden["out_parts_%d" % s], 7)
    assert open(path).read() == _text(texts, "parts_%s_00002.vtp" % name)


def test_vtp_takes_every_particle_below_the_count(tmp_path):
    import pic_io

    p = np.arange(21, dtype=float)
    path = str(tmp_path / "p.vtp")
    pic_io.write_vtp(path, "x", p, 10)
    text = open(path).read()
    assert 'NumberOfPoints="3"' in text


def test_diag_csv_matches_reference(golden, texts, tmp_path):
    import pic_io

    rows = golden["out_diag_rows"].reshape(2, -1)
    path = str(tmp_path / "runtime_diags.csv")
    for r, row in enumerate(rows):
        pic_io.diag_row(path, r == 0, ["e-", "O+"], int(row[0]), row[1], 0.0, row[2:14], row[14])
    got = open(path).read().splitlines()
    want = _text(texts, "runtime_diags.csv").splitlines()
    assert len(got) == len(want) == 3
    assert got[0] == want[0]
    for a, b in zip(got[1:], want[1:]):
        fa, fb = a.split(","), b.split(",")
        assert fa[:2] == fb[:2] and fa[3:] == fb[3:]  # column 2 is the wall clock
