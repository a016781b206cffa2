"""Host-side pieces of the far-wake surrogate that need no GPU: the
stratified operating points its training marches run at."""
import numpy as np
import pytest

from winds import surrogate as S


def test_operating_points_fill_every_cell_once():
    pts = S.operating_points(16, seed=17)
    assert len(pts) == 16
    cells = set()
    for t, y in pts:
        assert 4.0 <= t < 10.0 and 0.0 <= y < 30.0
        cells.add((int((t - 4.0) / 1.5), int(y / 7.5)))
    assert cells == {(i, j) for i in range(4) for j in range(4)}


def test_operating_points_seeded_and_box():
    a = S.operating_points(8, seed=3)
    assert a == S.operating_points(8, seed=3)
    assert a != S.operating_points(8, seed=4)
    b = S.operating_points(12, seed=3, tsr_range=(5.0, 6.0), yaw_range=(10.0, 20.0))
    assert all(5.0 <= t < 6.0 and 10.0 <= y < 20.0 for t, y in b)
    # 4 TSR bands x 3 yaw bands
    assert len({int((t - 5.0) / 0.25) for t, _ in b}) == 4
    assert len({int((y - 10.0) / (10.0 / 3)) for _, y in b}) == 3


def test_operating_points_need_a_multiple_of_four():
    with pytest.raises(ValueError):
        S.operating_points(6)
