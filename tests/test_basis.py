"""Basis layout of the product (aimd_amd.qc.basis) against the reference's layout rules and the oracle."""

import numpy as np
import pytest

from tests.conftest import HCL, METHANOL, WATER, bohr


def test_water_631gs_layout():
    """Reference tests/test_qc_integrals.py:75-91."""
    from aimd_amd.qc.basis import build_basis, cartesian_components
    s, x = bohr(WATER)
    b = build_basis(s, x, "6-31g*")
    assert (b.nao, b.nshell, b.natm) == (19, 10, 3)
    assert list(b.shell_l) == [0, 0, 0, 1, 1, 2, 0, 0, 0, 0]
    assert list(b.shell_atom) == [0] * 6 + [1, 1, 2, 2]
    np.testing.assert_array_equal(np.diff(b.shell_ao), (b.shell_l + 1) * (b.shell_l + 2) // 2)
    np.testing.assert_array_equal(b.atom_ao, [[0, 15], [15, 17], [17, 19]])
    np.testing.assert_array_equal(b.ao_lxyz[9:15], cartesian_components(2))
    assert b.ao_labels()[10] == "0 O dxy"
    assert b.shell_prim[1] - b.shell_prim[0] == 6
    with pytest.raises(ValueError):
        b.positions[0, 0] = 1.0


def test_basis_names_and_errors():
    from aimd_amd.qc.basis import build_basis
    from aimd_amd.qc.basis_data import normalize_basis_name
    assert normalize_basis_name("6-31G(d)") == "6-31g*"
    assert normalize_basis_name("cc-pVDZ") == "cc-pvdz"
    with pytest.raises(ValueError, match="Unknown basis set 'def2-svp'"):
        build_basis(["H"], [[0, 0, 0]], "def2-svp")
    with pytest.raises(ValueError, match="no data for element 'K'"):
        build_basis(["K"], [[0, 0, 0]], "sto-3g")
    with pytest.raises(ValueError, match="Unknown element"):
        build_basis(["Xx"], [[0, 0, 0]], "sto-3g")
    with pytest.raises(ValueError, match="expected 2 positions"):
        build_basis(["H", "H"], [[0, 0, 0]], "sto-3g")
    with pytest.raises(ValueError, match="finite"):
        build_basis(["H", "H"], [[0, 0, 0], [0, 0, np.nan]], "sto-3g")


def test_general_contractions_form_blocks_that_share_primitives():
    """Reference tests/test_qc_integrals.py:368-392: HCl/cc-pVDZ blocks, every primitive once."""
    from aimd_amd.qc.basis import build_basis
    s, x = bohr(HCL)
    b = build_basis(s, x, "cc-pvdz")
    cl = [B for B in range(b.nblock) if b.blk_atom[B] == 1]
    assert [(int(b.blk_l[B]), int(b.blk_ncon[B]), int(b.blk_prim[B + 1] - b.blk_prim[B]))
            for B in cl] == [(0, 3, 11), (0, 1, 1), (1, 2, 7), (1, 1, 1), (2, 1, 1)]
    assert b.nshell == 11 and b.nblock == 8 and b.nao == 24
    np.testing.assert_array_equal(b.blk_ao, b.shell_ao[b.blk_shell[:-1]])
    for B in range(b.nblock):
        s0, s1 = b.blk_shell[B], b.blk_shell[B + 1]
        assert len(set(b.shell_atom[s0:s1])) == 1 and len(set(b.shell_l[s0:s1])) == 1
        e = b.blk_exp[b.blk_prim[B]:b.blk_prim[B + 1]]
        assert len(set(e.tolist())) == len(e)        # each primitive of a block once
    p = build_basis(s, x, "6-31g*")
    assert p.nblock == p.nshell and np.all(p.blk_ncon == 1)


# reference tests/test_qc_integrals.py:395-412: zero-padded columns, an uncontracted
# shell repeating one exponent (merged into the block), a disjoint shell (own block)
GENERAL_DATA = {
    "O": [[0, [30.0, 0.2, 0.0], [6.0, 0.5, -0.2], [1.2, 0.4, 0.6]], [0, [1.2, 1.0]],
          [1, [5.0, 0.3, 0.1], [1.1, 0.6, -0.5], [0.3, 0.3, 0.9]],
          [2, [1.6, 0.7, 0.2], [0.5, 0.4, 0.8]], [2, [0.5, 1.0]], [2, [0.2, 1.0]]],
    "H": [[0, [4.0, 0.3, 0.1], [0.8, 0.7, 0.4], [0.2, 0.0, 0.8]], [1, [0.9, 1.0]]],
}
GENERAL_POS = np.array([[0.0, 0.1, -0.1], [1.5, 1.1, 0.2], [-1.4, 0.9, 0.4]])


def test_custom_general_contraction_blocks():
    from aimd_amd.qc.basis import basis_from_table
    b = basis_from_table(["O", "H", "H"], GENERAL_POS, GENERAL_DATA)
    o = [(int(b.blk_l[B]), int(b.blk_ncon[B])) for B in range(b.nblock) if b.blk_atom[B] == 0]
    assert o == [(0, 3), (1, 2), (2, 3), (2, 1)]
    assert int(b.blk_ncon.max()) == 3 and (b.nshell, b.nblock) == (15, 8)
    # the zero entry of the first O s column is dropped from that shell
    assert b.shell_prim[1] - b.shell_prim[0] == 3 and b.shell_prim[2] - b.shell_prim[1] == 2


@pytest.mark.parametrize("basis", ["sto-3g", "6-31g", "6-31g*", "6-31g**", "cc-pvdz"])
def test_product_basis_equals_oracle_restatement(oracle_lib, basis):
    from aimd_amd.qc.basis import build_basis
    s, x = bohr(METHANOL)
    b = build_basis(s, x, basis)
    o = oracle_lib.make_basis(s, x, basis)
    for name in ("shell_atom", "shell_l", "shell_prim", "shell_ao", "prim_exp",
                 "blk_atom", "blk_l", "blk_ncon", "blk_ao", "blk_prim", "blk_exp"):
        np.testing.assert_array_equal(getattr(b, name), getattr(o, name), err_msg=name)
    np.testing.assert_array_equal(b.blk_shell, o.blk_shell)
    # the product normalises with the closed-form primitive overlap (2 sqrt(ab)/(a+b))^(l+3/2),
    # the oracle with the reference's expression: the same numbers up to rounding
    np.testing.assert_array_max_ulp(b.prim_coef, o.prim_coef, maxulp=2)
    np.testing.assert_array_max_ulp(b.blk_coef, o.blk_coef, maxulp=2)


def test_with_positions_shares_arrays():
    from aimd_amd.qc.basis import build_basis
    s, x = bohr(WATER)
    b = build_basis(s, x, "6-31g*")
    m = b.with_positions(x + 0.1)
    assert np.shares_memory(m.prim_coef, b.prim_coef)
    np.testing.assert_array_equal(b.positions, x)
