"""The case matrix of the batched benchmark (host side only)."""
from iblbm_amd.ensemble import RE_NY, case_matrix
from iblbm_amd.init import grid_of, initialize


def test_case_matrix_72():
    cases = case_matrix()
    assert len(cases) == 72
    combos = {(c.Re, c.ibm_method, c.delta_type, c.collision_model) for c in cases}
    assert len(combos) == 72
    for c in cases:
        nx, ny, dx, dy = grid_of(c)
        assert ny == RE_NY[c.Re]
        assert nx == int(2.5 * (ny - 1)) + 1
        st = initialize(c, fields=False)
        assert abs(2 * st.lattice_r - (ny - 1) / 16) < 1e-9
        assert st.U0 is None and st.feq0 is None
    total = sum(grid_of(c)[0] * grid_of(c)[1] for c in cases)
    assert 2.9e8 < total < 3.1e8  # about 300 M lattice nodes


def test_shrunk_matrix_keeps_the_family():
    cases = case_matrix(shrink=8)
    for c in cases:
        nx, ny, _, _ = grid_of(c)
        assert ny == (RE_NY[c.Re] - 1) // 8 + 1
