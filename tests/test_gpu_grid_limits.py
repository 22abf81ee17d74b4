"""The FP32 and FP16-record movers index node records in 32 bits: a context on
a grid of 2^32 or more (ghost-inclusive, bricked in 3D) nodes is refused with
SP_ERR_UNSUPPORTED before any device memory is taken."""
import pytest

from helpers import cfg2d, cfg3d
from sp_hip import ParkerPusher
from sp_hip._abi import SpError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["f32", "f32h"])
@pytest.mark.parametrize("ndim", [2, 3])
def test_fp32_paths_refuse_grids_of_2_32_nodes(ndim, precision):
    # 2D: (65532 + 4)^2 = 2^32 nodes; 3D: (1628 + 4) = 204 bricks of 8 per axis,
    # 204^3 x 512 = 4.35e9 bricked nodes
    c = cfg2d(65532) if ndim == 2 else cfg3d(1628)
    print(f"grid limit: {ndim}D {precision}", flush=True)
    with pytest.raises(SpError) as ei:
        ParkerPusher(**dict(c, precision=precision, nptl_max=16))
    assert ei.value.code == -5 and "2^32" in str(ei.value)
