"""qmi_mlp_forward_noalloc, the forward pass of the overlap MLP on device
buffers that allocates nothing (its activations live in the library's fixed
per-device scratch array): through the C-ABI, two calls in a row -- the
second with a smaller batch, so stale scratch contents from the first call
would show -- give bitwise the outputs of qmi_mlp_forward on the same rows,
the call refuses a batch larger than the scratch, and it also runs on a
caller's stream."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net():
    from q_amd.mlp import MLP
    m = MLP(34, 128, seed=11)
    yield m
    m.close()


def test_noalloc_forward_matches_forward(net):
    from q_amd.engine import QmiError
    rng = np.random.default_rng(5)
    for batch in (1500, 97):
        x = rng.normal(size=(batch, 34)).astype(np.float32)
        ref = net.forward(x)
        dx = torch.from_numpy(x).to("cuda")
        dy = torch.full((batch,), -1.0, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        print(f"[gpu] noalloc forward, batch {batch}", flush=True)
        net.forward_device(dx.data_ptr(), batch, dy.data_ptr())
        torch.cuda.synchronize()
        y = dy.cpu().numpy()
        assert np.array_equal(y, ref), (batch, np.abs(y - ref).max())
    # on the caller's stream
    x = rng.normal(size=(64, 34)).astype(np.float32)
    ref = net.forward(x)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        dx = torch.from_numpy(x).to("cuda")
        dy = torch.zeros(64, dtype=torch.float32, device="cuda")
        net.forward_device(dx.data_ptr(), 64, dy.data_ptr(), stream=s.cuda_stream)
    s.synchronize()
    assert np.array_equal(dy.cpu().numpy(), ref)
    # beyond the scratch (2047 rows of 4 * 128 + 1 floats): refused, nothing launched
    big = torch.zeros((4096, 34), dtype=torch.float32, device="cuda")
    out = torch.zeros(4096, dtype=torch.float32, device="cuda")
    with pytest.raises(QmiError, match="exceeds the scratch"):
        net.forward_device(big.data_ptr(), 4096, out.data_ptr())
