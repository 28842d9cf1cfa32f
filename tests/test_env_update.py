"""The environment update entry point (ttg_env_update; updateEnv,
dmrg.cpp:491-509).

CPU: the index pairs the entry point uses for each direction equal the
pairs the oracle's operator* finds by tag (qtensor.cpp:540-555) on the
reference's layouts (psi = (l In, s In, r Out), W from the NN Heisenberg
MPO, TL = (l' In lvl1, w Out, l Out), TR = (r' Out lvl1, w In, r In)).

GPU: the updated environment equals the oracle's three-stage update: the
block list and offsets bit for bit, values to 1e-12 (FP64) or the reduced
forms' bounds; operands with another layout are rejected with an error."""
import numpy as np
import pytest

import oracle as O
from scenarios import mpo_bond
from ttgpu import lib as T
from ttgpu.flat import INWARD, LINK, OUTWARD, SITE, make_index, random_tensor

TOL = {T.FP64: 1e-12, T.FP32: 5e-6, T.BF16X3: 5e-5}


def sec(qs, ds):
    return [((q, 0), d) for q, d in zip(qs, ds)]


def env_case(direction, site=6, seed=7, scale=1):
    """psi_i, W_i and the environment on the side the update starts from."""
    rng = np.random.default_rng(seed)
    W, _ = mpo_bond("nn_chain", site)
    l = make_index(INWARD, "ID1Link%d" % site, LINK, 0,
                   sec([-4, -2, 0, 2, 4], [scale * d for d in (2, 5, 7, 4, 1)]))
    s = make_index(INWARD, "Site%d" % site, SITE, 0, sec([-1, 1], [1, 1]))
    r = make_index(OUTWARD, "ID1Link%d" % (site + 1), LINK, 0,
                   sec([-5, -3, -1, 1, 3, 5], [scale * d for d in (1, 4, 6, 6, 3, 2)]))
    psi = random_tensor([l, s, r], rng)
    if direction == T.MOVE_FROM_LEFT:
        lp = l.copy(); lp.level = 1
        env = random_tensor([lp, W.idx[0].dag(), l.dag()], rng)
    else:
        rp = r.copy(); rp.level = 1
        env = random_tensor([rp, W.idx[3].dag(), r.dag()], rng)
    return psi, W, env


def oracle_update(psi, W, env):
    """The oracle's stage pairs and result (dmrg.cpp:494-508)."""
    op, oe, ow = (O.OTensor.from_flat(t) for t in (psi, env, W))
    t1 = op.copy()
    t1.dag()
    t1.prime()
    pairs = [t1.pairs(oe)]
    t4 = t1 * oe
    pairs.append(t4.pairs(ow))
    t5 = t4 * ow
    pairs.append(t5.pairs(op))
    return pairs, (t5 * op).to_flat()


@pytest.mark.parametrize("direction", [T.MOVE_FROM_LEFT, T.MOVE_FROM_RIGHT])
def test_env_pairs_match_tag_pairing(direction):
    T.load()
    psi, W, env = env_case(direction)
    pairs, _ = oracle_update(psi, W, env)
    assert T.env_pairs(direction) == pairs


@pytest.mark.gpu
@pytest.mark.parametrize("direction", [T.MOVE_FROM_LEFT, T.MOVE_FROM_RIGHT])
@pytest.mark.parametrize("prec", [T.FP64, T.FP32, T.BF16X3])
def test_env_update_matches_oracle(direction, prec):
    psi, W, env = env_case(direction, scale=9)
    _, ref = oracle_update(psi, W, env)
    dp, dw, de = (T.DTensor.from_flat(t) for t in (psi, W, env))
    got = T.env_update(dp, de, dw, direction, prec)
    qi, off = got.meta()
    assert np.array_equal(qi, ref.block_qi)
    assert np.array_equal(off, ref.block_offsets())
    err = np.linalg.norm(got.data() - ref.data) / np.linalg.norm(ref.data)
    assert err <= TOL[prec], err
    assert [ix.arrow for ix in got.idx] == [ix.arrow for ix in ref.idx]
    assert [ix.level for ix in got.idx] == [ix.level for ix in ref.idx]
    # the inputs are unchanged (the entry point reads psi through a dag view)
    assert np.array_equal(dp.data(), psi.data)
    assert np.array_equal(de.data(), env.data)


@pytest.mark.gpu
def test_env_update_rejects_another_layout():
    psi, W, env = env_case(T.MOVE_FROM_LEFT)
    dp, dw, de = (T.DTensor.from_flat(t) for t in (psi, W, env))
    with pytest.raises(RuntimeError, match="layout"):
        T.env_update(dp, de, dw, T.MOVE_FROM_RIGHT)   # TL given, TR expected
    # the library keeps working
    got = T.env_update(dp, de, dw, T.MOVE_FROM_LEFT)
    assert got.info()[1] > 0
