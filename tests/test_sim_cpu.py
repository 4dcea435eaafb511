"""READINI's `bubbin` reading (include/eb_sim.h, host code): the reference's own input file
and list-directed variants of it."""
import os

import pytest

from tests.conftest import GOLDEN


@pytest.fixture(scope="module")
def sim():
    from electricbubbles_amd import sim as _s
    return _s


def test_reads_the_reference_bubbin(sim):
    """tests/golden/bubbin_4circles is the first run description of the reference's bubbin."""
    p = sim.read_bubbin(os.path.join(GOLDEN, "bubbin_4circles")).to_dict()
    assert p["k"] == 4 and p["nd"] == [128] * 4 and p["nplot"] == 10
    assert p["dt"] == 0.01 and p["tfinal"] == 0.5
    assert (p["cap"], p["bondx"], p["bondy"], p["e0"]) == (0.0, 0.0, 0.0, 0.64)
    assert p["cx"] == [-1.5, -1.5, 1.5, 1.5] and p["cy"] == [1.5, -1.5, 1.5, -1.5]
    assert p["ai"] == [1.0] * 4 and p["bi"] == [1.0] * 4 and p["ncyc"] == [0] * 4


def test_list_directed_variants(sim, tmp_path):
    """Values split over records, commas, D exponents, trailing values skipped."""
    f = tmp_path / "bubbin"
    f.write_text("2 extra\n256,\n512\n1.d-3, 2.5D0 20\n0.1 0 0 1.44\n1.1 0.9 0.2 0. 0. 0\n1. .1 0. 3. 0. 3\n")
    p = sim.read_bubbin(str(f)).to_dict()
    assert p["k"] == 2 and p["nd"] == [256, 512]
    assert p["dt"] == 1e-3 and p["tfinal"] == 2.5 and p["nplot"] == 20
    assert p["cap"] == 0.1 and p["e0"] == 1.44
    assert p["ai"] == [1.1, 1.0] and p["bi"] == [0.9, 0.1] and p["ncyc"] == [0, 3] and p["cx"] == [0.0, 3.0]


def test_incomplete_file_is_an_error(sim, tmp_path):
    import electricbubbles_amd as eb
    f = tmp_path / "bubbin"
    f.write_text("3\n128 128 128\n.01 .5 10\n0 0 0 .64\n1. 1. 0. 0. 0. 0\n")
    with pytest.raises(eb.EBError):
        sim.read_bubbin(str(f))
