"""The device supercell builder (tbg_supercell_hops) against the loop
restatement of SuperCell.get_hop, bit for bit: graphene periodic and open
(and against the graphene generator at benchmark-like size), a 3-D
three-orbital cell with long hops, mixed boundaries and vacancies."""
import math

import numpy as np
import pytest

import oracle
import tbplas_amd as tb

pytestmark = pytest.mark.gpu

LAT = np.array([[0.246, 0, 0], [0.246 * math.cos(math.pi / 3), 0.246 * math.sin(math.pi / 3), 0],
                [0, 0, 1.0]])
FRAC = np.array([[0.0, 0.0, 0.0], [1 / 3., 1 / 3., 0.0]])
IND = np.array([(0, 0, 0, 0, 1), (1, 0, 0, 1, 0), (0, 1, 0, 1, 0)])
ENG = np.full(3, -2.7 + 0j)


def _same(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("dims,pbc", [((6, 5, 1), (True, True, False)),
                                      ((7, 4, 1), (True, False, False)),
                                      ((3, 3, 1), (False, False, False))])
def test_graphene_cell(gpu, dims, pbc):
    got = tb.supercell_hops(LAT, FRAC, IND, ENG, dims, pbc)
    ref = oracle.orc_supercell_hops(LAT, FRAC, IND, ENG, dims, pbc)
    _same(got, ref)


def test_graphene_cell_matches_generator_large(gpu):
    """2 x 128 x 128 orbitals: the general builder and the graphene
    generator give the same arrays."""
    pos, hi, hj, hv, dr = tb.supercell_hops(LAT, FRAC, IND, ENG, (128, 128, 1),
                                            (True, True, False))
    gp, gi, gj, gv, gd = tb.graphene_hops(128, 128)
    _same((pos, hi, hj, hv, dr), (gp, gi, gj, gv, gd))


def _odd_cell():
    lat = np.array([[0.31, 0.02, 0.0], [-0.05, 0.27, 0.01], [0.0, 0.03, 0.42]])
    frac = np.array([[0.0, 0.0, 0.0], [0.5, 0.1, 0.2], [0.2, 0.6, 0.7]])
    ind = np.array([(0, 0, 0, 0, 1), (0, 0, 0, 1, 2), (1, 0, 0, 0, 0), (2, -1, 0, 2, 1),
                    (0, 1, 1, 1, 0), (-1, 2, 0, 0, 2), (0, 0, 1, 2, 2), (1, 1, -1, 1, 2)])
    eng = np.array([-1.0, 0.5 + 0.2j, -0.3, 0.1j, 0.7, -0.05 + 0.01j, 0.2, -0.4])
    return lat, frac, ind, eng


@pytest.mark.parametrize("pbc", [(True, True, True), (True, False, True), (False, False, False)])
@pytest.mark.parametrize("vac", [[], [(0, 1, 1, 2), (2, 3, 0, 0), (1, 0, 1, 1), (0, 0, 0, 0)]])
def test_three_orbital_cell(gpu, pbc, vac):
    lat, frac, ind, eng = _odd_cell()
    got = tb.supercell_hops(lat, frac, ind, eng, (3, 4, 2), pbc, vac)
    ref = oracle.orc_supercell_hops(lat, frac, ind, eng, (3, 4, 2), pbc, vac)
    _same(got, ref)


def test_bad_arguments(gpu):
    lat, frac, ind, eng = _odd_cell()
    bad = ind.copy()
    bad[0, 4] = 7
    with pytest.raises(tb._lib.TBGError):
        tb.supercell_hops(lat, frac, bad, eng, (2, 2, 2))
    with pytest.raises(tb._lib.TBGError):
        tb.supercell_hops(lat, frac, ind, eng, (0, 2, 2))
