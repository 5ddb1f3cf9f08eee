"""The allocation-free loss and gradient of the radius network
(gts_mlp_loss_grad_noalloc, device buffers, a fixed per-GPU array for the
activations) against the allocating entry point (gts_mlp_loss_grad) on the
same inputs: gradients identical (the same kernels, per-parameter sums in a
fixed order), loss within 1e-6 relative (its per-block partial sums are added
atomically); two calls in a row, the second with a smaller batch; a batch
above the array's capacity is refused.
"""
import numpy as np
import pytest

from gts_amd import GTS_E_ARG, MLP_NOALLOC_MAX_NB, DeviceArray, GTSError, GTSIndex, datagen

pytestmark = pytest.mark.gpu


def _bf16_round(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def test_noalloc_loss_grad_matches_allocating_entry_point():
    rng = np.random.default_rng(21)
    idx = GTSIndex(datagen.Dataset(2, vectors=rng.random((200, 4)).astype(np.float32)))
    shapes = [(64, 16), (64,), (64, 64), (64,), (1, 64), (1,)]
    w = np.concatenate([_bf16_round(rng.standard_normal(s) * 0.3).ravel() for s in shapes]).astype(np.float32)
    grad_d = idx.device_array(w.size, np.float32)
    loss_d = idx.device_array(1, np.float32)
    w_d = DeviceArray.from_host(idx, w)
    for nb in (MLP_NOALLOC_MAX_NB, 700):    # the second call smaller: stale rows must not leak in
        X = rng.standard_normal((nb, 16)).astype(np.float32)
        t = rng.standard_normal(nb).astype(np.float32)
        ref_loss, ref_g = idx.mlp_loss_grad(X, t, w)
        X_d, t_d = DeviceArray.from_host(idx, X), DeviceArray.from_host(idx, t)
        idx.mlp_loss_grad_noalloc(X_d, t_d, w_d, grad_d, loss_d)
        g = grad_d.to_host()
        loss = float(loss_d.to_host()[0])
        print("nb=%d loss %.7g (allocating %.7g), max |dg| %.3g" % (nb, loss, ref_loss,
              np.abs(g - ref_g).max()), flush=True)
        assert np.array_equal(g, ref_g)
        assert abs(loss - ref_loss) <= 1e-6 * max(1.0, abs(ref_loss))
    big = MLP_NOALLOC_MAX_NB + 1
    X_d = DeviceArray.from_host(idx, np.zeros((big, 16), np.float32))
    t_d = DeviceArray.from_host(idx, np.zeros(big, np.float32))
    with pytest.raises(GTSError) as e:
        idx.mlp_loss_grad_noalloc(X_d, t_d, w_d, grad_d, loss_d)
    assert e.value.code == GTS_E_ARG
