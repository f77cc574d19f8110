"""Rate families (src/compiled_families.jl:154-300): argument validation on
the host, following test/test_compiled_families.jl:1-110; Ritz ordering of
the spectrum driver (src/krylov.jl:893-898)."""
import numpy as np
import pytest

import pimi
from pimi.families import _family_rates
from pimi.spectral import _ritz_order

SP = pimi.spin_matrices(2)


def _prototype():
    return pimi.PIModel(pimi.PIBasis(2, 2), (pimi.LocalHamiltonian(SP.jx, rate=0.17),
                                             pimi.LocalJump(SP.jm, rate=0.8), pimi.LocalJump(SP.jp, rate=0.2)))


def test_family_selection_and_defaults():
    fam = pimi.compile_family(_prototype())
    assert fam.rate_indices == (0, 1, 2)
    assert fam.default_rates == (0.17, 0.8, 0.2)
    assert fam.estimates["geometry_reused"] and fam.estimates["dimension"] == 10
    one = pimi.compile_family(_prototype(), rate_indices=(2,))
    assert _family_rates(one, 0.31) == (0.31,)
    with pytest.raises(ValueError):
        _family_rates(fam, 0.31)
    with pytest.raises(ValueError):
        _family_rates(fam, (1.0, 2.0))
    with pytest.raises(ValueError):
        _family_rates(fam, (1.0, 2.0, np.inf))
    with pytest.raises(ValueError):
        _family_rates(fam, (1.0, 2.0, 1j))


def test_family_rejects_bad_indices_and_rates():
    with pytest.raises(ValueError):
        pimi.compile_family(_prototype(), rate_indices=[])
    with pytest.raises(ValueError):
        pimi.compile_family(_prototype(), rate_indices=(1, 1))
    with pytest.raises(IndexError):
        pimi.compile_family(_prototype(), rate_indices=(3,))
    b = pimi.PIBasis(2, 2)
    with pytest.raises(ValueError):
        pimi.compile_family(pimi.PIModel(b, (pimi.LocalJump(SP.jm, rate=np.inf),)))
    with pytest.raises(ValueError):
        pimi.compile_family(pimi.PIModel(b, (pimi.LocalJump(SP.jm, rate=1j),)))


def test_ritz_order():
    v = np.array([-1 + 0j, 0.5 + 2j, -0.1 + 0j, 3j])
    np.testing.assert_array_equal(_ritz_order(v, "LR"), [1, 3, 2, 0])
    np.testing.assert_array_equal(_ritz_order(v, "LM"), [3, 1, 0, 2])
    np.testing.assert_array_equal(_ritz_order(v, "SM"), [2, 0, 1, 3])
    with pytest.raises(ValueError):
        _ritz_order(v, "SR")
