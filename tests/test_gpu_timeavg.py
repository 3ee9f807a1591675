"""Time averaging (MomentTimeAvg, moments/MomentTimeAvg.h:19-80) of the HIP
core vs the CPU oracle, through the C ABI.

The add is fused into the next step's kernel (which reads the state anyway);
a read of the averages, a change of the state or a network step adds a
pending state with a pass of its own.  FP64 literal must equal the oracle's
averages bit for bit; in the fast arithmetic the fused add must equal the
mean of the lbm_get_macroscopic snapshots of the same states bit for bit.
"""
import numpy as np
import pytest

from helpers import OracleLattice, bc_nodes, random_velocity_field
from lb_amd import Lattice, abi, make_desc
from oracle import mesher_oracle as mo

pytestmark = pytest.mark.gpu


def _channel_pair(**kw):
    flags, bcd = mo.channel(24, 12, 10)
    bc = bc_nodes(bcd)
    print(f"[timeavg channel] kw={kw}", flush=True)
    g = Lattice(make_desc(flags, bc, **kw))
    o = OracleLattice(make_desc(flags, bc, **kw))
    rng = np.random.default_rng(12)
    solid = flags == 0
    init = (None, np.where(solid, 0.0, 0.04 + 0.004 * rng.standard_normal(flags.shape)),
            np.where(solid, 0.0, 0.004 * rng.standard_normal(flags.shape)), None)
    g.init_equilibrium(*init)
    o.init_equilibrium(*init)
    return flags, g, o


def _assert_same_avg(flags, g, o, tag):
    (ga, gn), (oa, on) = g.moment_avg(), o.moment_avg()
    fluid = flags != 0
    err = max(np.abs(a[fluid] - b[fluid]).max() for a, b in zip(ga, oa))
    print(f"[{tag}] count {gn}/{on} max|d avg| {err:.3e}", flush=True)
    assert gn == on
    for a, b in zip(ga, oa):
        assert np.array_equal(a[fluid], b[fluid])


@pytest.mark.parametrize("coll", [abi.COLLISION_CUM2015, abi.COLLISION_COMBINED], ids=["cum2015", "combined"])
def test_channel_fp64_literal_average_equals_oracle(gpu, coll):
    flags, g, o = _channel_pair(collision=coll, closure=abi.CLOSURE_SMAGORINSKY, c_les=0.17,
                                storage=abi.STORAGE_F64, arith=abi.ARITH_LITERAL, compat=abi.COMPAT_REFERENCE,
                                tau=0.53, nu=0.01)
    # the initial state (first-step fields), then a window of counted steps
    g.moment_add()
    o.moment_add()
    g.step(4)
    o.step(4)
    g.moment_add()
    o.moment_add()
    g.step_moment_add(7)
    o.step_moment_add(7)
    _assert_same_avg(flags, g, o, f"channel coll={coll} after 7")
    # uncounted steps, then a second window; the sums continue
    g.step(3)
    o.step(3)
    g.step_moment_add(5)
    o.step_moment_add(5)
    _assert_same_avg(flags, g, o, f"channel coll={coll} after 5 more")
    # reset, and a pending add resolved by a state change
    g.moment_avg_reset()
    g.step_moment_add(2)
    g.set_populations(g.populations())
    avg, n = g.moment_avg()
    assert n == 2


@pytest.mark.parametrize("storage", [abi.STORAGE_F32, abi.STORAGE_F16DEV_SR], ids=["f32", "f16dev_sr"])
def test_fast_fused_add_equals_snapshot_mean(gpu, storage):
    n = 32
    flags = np.ones((n, n, n), np.int32)
    g = Lattice(make_desc(flags, [], collision=abi.COLLISION_CUM2017, closure=abi.CLOSURE_SMAGORINSKY, c_les=0.17,
                          storage=storage, arith=abi.ARITH_FAST, compat=abi.COMPAT_REFERENCE, periodic=(1, 1, 1),
                          tau=0.5015))
    u = random_velocity_field(flags.shape, 0.04, seed=40)
    g.init_equilibrium(None, *u)
    acc = [np.zeros(flags.shape) for _ in range(4)]
    g.moment_add()
    for a, m in zip(acc, g.macroscopic()):
        a += m
    for k in range(6):
        g.step_moment_add(1)
        for a, m in zip(acc, g.macroscopic()):
            a += m
    avg, cnt = g.moment_avg()
    assert cnt == 7
    for a, b in zip(avg, acc):
        assert np.array_equal(a, b / 7)


def test_fused_add_over_many_steps_equals_snapshot_mean(gpu):
    """Seven counted steps in one call: every add but the last is fused."""
    n = 24
    flags = np.ones((n, n, n), np.int32)
    kw = dict(collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY, c_les=0.17,
              storage=abi.STORAGE_F32, arith=abi.ARITH_FAST, compat=0, periodic=(1, 1, 1), tau=0.51)
    u = random_velocity_field(flags.shape, 0.04, seed=41)
    a = Lattice(make_desc(flags, [], **kw))
    b = Lattice(make_desc(flags, [], **kw))
    a.init_equilibrium(None, *u)
    b.init_equilibrium(None, *u)
    a.step(2)
    b.step(2)
    a.step_moment_add(7)
    acc = [np.zeros(flags.shape) for _ in range(4)]
    for k in range(7):
        b.step(1)
        for s, m in zip(acc, b.macroscopic()):
            s += m
    avg, cnt = a.moment_avg()
    assert cnt == 7
    assert np.array_equal(a.populations(), b.populations())
    for x, y in zip(avg, acc):
        assert np.array_equal(x, y / 7)


def test_average_with_the_learned_closure(gpu):
    """Network steps add a pending state with their own pass; held-coefficient
    steps fuse it."""
    n = 16
    flags = np.ones((n, n, n), np.int32)
    g = Lattice(make_desc(flags, [], collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_MLP, storage=abi.STORAGE_F32,
                          arith=abi.ARITH_FAST, compat=0, periodic=(1, 1, 1), tau=0.503, c_les=0.17))
    rng = np.random.default_rng(7)
    p = np.zeros(abi.MLP_PARAMS, np.float32)
    p[:768] = rng.standard_normal(768) / np.sqrt(12.0)
    p[832:4928] = rng.standard_normal(4096) / 8.0
    p[4992:5056] = 0.05 * rng.standard_normal(64)
    norm = np.stack([np.zeros(12), np.ones(12), np.full(12, -1e30), np.full(12, 1e30)]).astype(np.float32)
    g.set_closure(p, norm)
    g.set_closure_refresh(3)
    u = random_velocity_field(flags.shape, 0.04, seed=42)
    g.init_equilibrium(None, *u)
    acc = [np.zeros(flags.shape) for _ in range(4)]
    for k in range(7):
        g.step_moment_add(1)
        for s, m in zip(acc, g.macroscopic()):
            s += m
    avg, cnt = g.moment_avg()
    assert cnt == 7
    for x, y in zip(avg, acc):
        assert np.array_equal(x, y / 7)
