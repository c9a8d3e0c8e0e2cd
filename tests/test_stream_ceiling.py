"""The streaming-ceiling pass (cbm_stream_ceiling): the bytes it reports and
its argument checks."""
import pytest

from cerebellum_simulation_amd import CbmError, stream_ceiling

pytestmark = pytest.mark.gpu


def test_stream_ceiling_reports_bytes_and_time(gpu):
    print("[case] stream ceiling 2^16 + 100 cells", flush=True)
    cells = (1 << 16) + 100  # a ragged last block
    us, nbytes = stream_ceiling(cells, read_words=7, write_words=6, launches=5)
    assert nbytes == cells * (7 * 4 + 16 + 6 * 4)
    assert 0 < us < 1e5


@pytest.mark.parametrize("args", [(0, 7, 6), (1024, 0, 0), (1024, 9, 6), (1024, 3, 4)])
def test_stream_ceiling_rejects_bad_arguments(gpu, args):
    with pytest.raises(CbmError, match="bad argument"):
        stream_ceiling(*args, launches=1)
