"""Pins the oracle's model tables against closed forms computed here from the
model arrays (MpdRdmeSolver.cu:408-421 transition table, :697-774
propensity tables), including the reference quirks: hetero second-order
entries are assigned (the last reaction on a pair wins) and self reactions
enter with a factor 2."""
import numpy as np

import oracle
from helpers import SPACING, TAU, spec


def f32(x):
    return np.float32(x)


def test_transition_table_closed_form():
    DF = np.array([[[1e-12, 3e-13], [0.0, 2.5e-12]], [[7e-14, 0.0], [4.9e-11, 1e-16]]])
    s = spec((4, 4, 4), nS=2, nST=2, DF=DF)
    o = oracle.Oracle(s)
    want = (DF * TAU / (SPACING * SPACING)).astype(np.float32).reshape(-1)
    assert o.T.view(np.uint32).tolist() == want.view(np.uint32).tolist()


def test_transition_table_rejects_q_above_half():
    DF = np.full((1, 1, 1), 0.51 * SPACING ** 2 / TAU)
    import pytest

    with pytest.raises(ValueError):
        oracle.Oracle(spec((4, 4, 4), DF=DF))


def test_propensity_tables_closed_forms():
    X, Y, Z = 6, 5, 4
    N = X * Y * Z
    # 0: zero order -> species 2; 1: first order A -> B; 2: A + B -> C;
    # 3: self 2C -> A;  4: A + B -> A (same pair as 2: overwrites it)
    rx = [(0, 3.0, {2: 1}, []),
          (1, 7.0, {0: -1, 1: 1}, [0]),
          (2, 1e-3, {0: -1, 1: -1, 2: 1}, [0, 1]),
          (3, 2e-3, {2: -2, 0: 1}, [2]),
          (2, 5e-3, {1: -1}, [0, 1])]
    s = spec((X, Y, Z), nS=3, nST=2, reactions=rx)
    s.RL[:, 1] = 0
    s.RL[1, 1] = 1
    o = oracle.Oracle(s)
    k = np.array([r[1] for r in rx])
    rate = [f32(k[0] * TAU / N), f32(k[1] * TAU), f32(k[2] * TAU * X * Y * Z), f32(k[3] * TAU * X * Y * Z),
            f32(k[4] * TAU * X * Y * Z)]
    assert o.rate[:5].tolist() == rate
    scale = float(np.float32(N))
    qp0 = o.qp0.reshape(2)
    qp1 = o.qp1.reshape(2, 3)
    qp2 = o.qp2.reshape(2, 3, 3)
    assert qp0[0] == f32(0.0 + k[0] * TAU / scale) and qp0[1] == 0.0
    assert qp1[0, 0] == f32(k[1] * TAU) and qp1[1, 0] == f32(k[1] * TAU)
    assert qp1[0, 1] == 0.0 and qp1[0, 2] == 0.0
    # the later reaction on the pair (A, B) overwrites the earlier one
    assert qp2[0, 0, 1] == f32(k[4] * TAU * scale) == qp2[0, 1, 0]
    # self reaction: factor 2
    assert qp2[0, 2, 2] == f32(0.0 + k[3] * TAU * scale * 2)
    assert np.count_nonzero(qp2[1]) == 0
    # reaction tables (buildModel): orders, D1/D2 = species+1, S reaction-major
    assert o.order[:5].tolist() == [1, 2, 3, 4, 3]
    assert o.D1[:5].tolist() == [0, 1, 1, 3, 1]
    assert o.D2[:5].tolist() == [0, 0, 2, 0, 2]
    assert o.S_rm.reshape(5, 3)[3].tolist() == [1, 0, -2]


def test_invalid_dependencies_rejected():
    import pytest

    # first-order reaction with two dependencies
    s = spec((4, 4, 4), nS=2, reactions=[(1, 1.0, {0: -1}, [0, 1])])
    with pytest.raises(ValueError):
        oracle.Oracle(s)
    # zeroth-order with a dependency
    s = spec((4, 4, 4), nS=2, reactions=[(0, 1.0, {0: 1}, [0])])
    with pytest.raises(ValueError):
        oracle.Oracle(s)
