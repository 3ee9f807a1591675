"""lbm_step_diagnostics_noalloc against lbm_step_diagnostics, through the C
ABI: two identical lattices, one stepped by each entry point, give the same
seven values (to FP64 summation rounding: the one sums its tile records on
the host, the other on the device) and the same state bit for bit; then a
second, smaller lattice through the same shared device array.  A lattice
with more tiles than the array holds is refused before any step."""
import numpy as np
import pytest

from helpers import bc_nodes
from lb_amd import Lattice, LbmError, abi, make_desc
from oracle import mesher_oracle as mo

pytestmark = pytest.mark.gpu


def _pair(nx, ny, nz, storage):
    flags, bcd = mo.channel(nx, ny, nz)
    bc = bc_nodes(bcd)
    kw = dict(collision=abi.COLLISION_CUM2015, closure=abi.CLOSURE_SMAGORINSKY, c_les=0.17, storage=storage,
              arith=abi.ARITH_FAST, compat=abi.COMPAT_REFERENCE, tau=0.53, nu=0.01)
    rng = np.random.default_rng(nx)
    solid = flags == 0
    ux = np.where(solid, 0.0, 0.04 + 0.004 * rng.standard_normal(flags.shape))
    uy = np.where(solid, 0.0, 0.004 * rng.standard_normal(flags.shape))
    out = []
    for _ in range(2):
        L = Lattice(make_desc(flags, bc, **kw))
        L.init_equilibrium(None, ux, uy, None)
        L.step(9)
        out.append(L)
    return out


def _close(a, b):
    return abs(a - b) <= 1e-12 * max(abs(a), abs(b), 1e-300) + 1e-300


@pytest.mark.parametrize("storage", [abi.STORAGE_F64, abi.STORAGE_F32], ids=["f64", "f32"])
def test_noalloc_diagnostics_equal_the_allocating_entry_point(gpu, storage):
    for dims in ((96, 48, 40), (40, 20, 16)):      # the second call is on a smaller lattice
        a, b = _pair(*dims, storage)
        for rep in range(2):
            ma, ja, fa = a.step_diagnostics()
            mb, jb, fb = b.step_diagnostics_noalloc()
            print(f"[noalloc {dims} {abi.STORAGE_NAMES[storage]} call {rep}] mass {ma:.15e} / {mb:.15e}, "
                  f"force x {fa[0]:.6e} / {fb[0]:.6e}", flush=True)
            assert _close(ma, mb)
            for u, v in zip(np.concatenate([ja, fa]), np.concatenate([jb, fb])):
                assert _close(u, v)
            assert abs(fa[0]) > 0          # the obstacle and walls take momentum
        assert np.array_equal(a.populations(), b.populations())
        assert a.steps() == b.steps()
        a.close()
        b.close()


def test_noalloc_refuses_a_lattice_beyond_its_array(gpu):
    """1024 x 256 x 257 cells, FP16 deviations: 263 168 tiles > 262 144."""
    flags = np.ones((257, 256, 1024), np.int32)
    L = Lattice(make_desc(flags, [], collision=abi.COLLISION_CUM2015, storage=abi.STORAGE_F16DEV,
                          arith=abi.ARITH_FAST, compat=0, periodic=(1, 1, 1), tau=0.6))
    L.init_equilibrium(None, None, None, None)
    with pytest.raises(LbmError) as ei:
        L.step_diagnostics_noalloc()
    assert "LBM_E_UNSUPPORTED" in str(ei.value)
    assert L.steps() == 0
    L.close()
