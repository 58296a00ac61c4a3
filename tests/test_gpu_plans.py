"""Default plans of the black-cell RBGS passes (tile shape and z-chunk length a context
starts with). Plans never change results (test_gpu_parity.py steps every tile shape and
chunk length bit-exactly); these tests pin the choices that make the default fast."""
import pytest

from conftest import sf

pytestmark = pytest.mark.gpu


def default_plan(shape):
    desc = sf.scenes.plume_desc(max(shape), 2, 2, shape=shape)
    ctx = sf.Context(desc)
    try:
        return ctx.smoother_tiling()
    finally:
        ctx.destroy()


@pytest.mark.parametrize("shape,tile", [
    ((64, 64, 64), sf.abi.TILE_64X16),      # 128-cell tiles would leave half their x positions idle
    ((192, 192, 192), sf.abi.TILE_64X16),   # ... or a quarter of the last tile column
    ((32, 48, 40), sf.abi.TILE_64X16),
    ((96, 96, 96), sf.abi.TILE_128X16),     # both shapes leave 16 pairs idle
    ((128, 160, 128), sf.abi.TILE_128X16),
    ((256, 256, 256), sf.abi.TILE_128X16),
])
def test_default_tile_fits_the_width(shape, tile):
    assert default_plan(shape)[0] == tile


def test_default_chunks_balance_the_cus():
    """256^3: 16-plane chunks (512 workgroups, two per CU). On a 256-CU device the reference
    benchmark's 128x160x128 grid takes 6-plane chunks (220 workgroups: none of the CUs runs
    two while others run one, as 4-plane chunks, 320 workgroups, would) and 512^3 one
    workgroup per CU with 256-plane chunks."""
    assert default_plan((256, 256, 256)) == (sf.abi.TILE_128X16, 16)
    if sf.hip.compute_units() != 256:
        pytest.skip("chunk choices below are stated for 256 compute units")
    assert default_plan((128, 160, 128)) == (sf.abi.TILE_128X16, 6)
    assert default_plan((512, 512, 512)) == (sf.abi.TILE_128X16, 256)
