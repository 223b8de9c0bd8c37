"""LeMonADE BFM files: the reader and writer behind bfm_read_bfm_file /
bfm_write_bfm_file, against a hand-decoded fixture in the format as the
reference's own parser reads it (src/pscbfm/mainColorBFM.cpp:135-276), by
write/read fixed points, and their error reports.  Host only."""
from pathlib import Path

import numpy as np
import pytest

from pscbfm_amd import BfmError, UpdaterGPUScBFM_AB_Type, systems, updater_from_system

GOLDEN = Path(__file__).resolve().parent / "golden"


def read(path):
    u = UpdaterGPUScBFM_AB_Type()
    u.readBfmFile(str(path))
    return u


def test_fixture_decodes_to_the_last_conformation(tmp_path):
    # !mcs=12 holds two chains: (1,2,3) +(2,1,0) +(-2,0,0) and (9,8,7) +(0,0,2) +(0,2,0)
    u = read(GOLDEN / "two_chains.bfm")
    assert u.getAge() == 12
    want = np.array([[1, 2, 3], [3, 3, 3], [1, 3, 3], [9, 8, 7], [9, 8, 9], [9, 10, 9]], dtype=np.int32)
    assert np.array_equal(u.getMonomerPositions(6), want)
    # written back: box, periodicity, the extra bond 1-6 (not a chain bond),
    # attributes as ranges, both chains
    out = tmp_path / "back.bfm"
    u.writeBfmFile(str(out))
    text = out.read_bytes().decode("latin-1")
    assert "!box_x=16\n!box_y=16\n!box_z=16\n" in text
    assert "!periodic_x=1\n!periodic_y=1\n!periodic_z=0\n" in text
    assert "!bonds\n1 6\n\n" in text
    assert "!attributes\n1-3:1\n4-6:2\n\n" in text
    assert "!mcs=12\n" in text


def test_bond_vector_components_are_read_in_order(tmp_path):
    # a set holding only (2,1,0) and (-2,0,0): x and y must not be swapped or repeated
    u = read(GOLDEN / "two_chains.bfm")
    xyz = u.getMonomerPositions(6)
    assert tuple(xyz[1] - xyz[0]) == (2, 1, 0)
    assert tuple(xyz[2] - xyz[1]) == (-2, 0, 0)


@pytest.mark.parametrize("make", [
    lambda: systems.linear_melt(32, 64, 16, seed=3),
    lambda: systems.diamond_network(64, (2, 2, 8), 32, seed=4),
])
def test_write_read_write_is_a_fixed_point(tmp_path, make):
    s = make()
    n = len(s["xyz"])
    u = updater_from_system(s, seed=1, validate_only=True, age=321)
    a = tmp_path / "a.bfm"
    u.writeBfmFile(str(a))
    v = read(a)
    assert v.getAge() == 321
    assert np.array_equal(v.getMonomerPositions(n), np.asarray(s["xyz"]).reshape(-1, 3))
    v.validate()                        # bonds and bond set survived: the system is still legal
    b = tmp_path / "b.bfm"
    v.writeBfmFile(str(b))
    assert a.read_bytes() == b.read_bytes()
    # colours depend on the bonds only
    assert np.array_equal(u.getColors(n)[0], v.getColors(n)[0])


def test_appended_conformations_and_last_one_wins(tmp_path):
    s = systems.linear_melt(32, 32, 8, seed=5)
    n = len(s["xyz"])
    u = updater_from_system(s, seed=1, validate_only=True)
    f = tmp_path / "traj.bfm"
    u.writeBfmFile(str(f))
    u.setAge(100)
    u.writeBfmFile(str(f), append=True)
    text = f.read_bytes().decode("latin-1")
    assert text.count("!mcs=") == 2 and "!mcs=100\n" in text
    assert read(f).getAge() == 100


@pytest.mark.parametrize("body,needle", [
    ("!number_of_monomers=2\n!set_of_bondvectors\n2 0 0:17\n\n!mcs=0\n0 0 0 \x12\n", "bond byte 18"),
    ("!number_of_monomers=3\n!set_of_bondvectors\n2 0 0:17\n\n!mcs=0\n0 0 0 \x11\n", "holds 2 monomers"),
    ("!number_of_monomers=1\n!set_of_bondvectors\n2 0 0:17\n", "no conformation"),
    ("!number_of_monomers=2\n!bonds\n1 3\n", "out of range"),
    ("!number_of_monomers=2\n!set_of_bondvectors\n5 0 0:17\n", "[-4, 3]"),
])
def test_malformed_files_are_reported(tmp_path, body, needle):
    f = tmp_path / "bad.bfm"
    f.write_bytes(body.encode("latin-1"))
    with pytest.raises(BfmError) as e:
        read(f)
    assert needle in str(e.value)
