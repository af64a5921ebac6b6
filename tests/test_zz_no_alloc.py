"""mrtxGuideNetworkTrainNoAlloc (the training step whose workspace is a fixed
device array, for callers that cannot allocate) against the allocating
mrtxGuideNetworkTrain on the same records, through the C ABI.

Two networks start from the same seed.  Each step, one trains through
mrtxGuideNetworkTrain from host records, the other through the no-alloc
entry point from the same records in device memory; parameters, both Adam
moments and the step counter must be bit-identical, and the loss (a float
atomic sum over waves in both) equal to 1e-6.  The second
call uses fewer records than the first, so any state the fixed array keeps
from the earlier call would show.  Sizes beyond the array are rejected.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _records(n, seed):
    from miniraytrace_amd.guide import make_records
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3)).astype(np.float32)
    nrm, wo, wi = _unit(rng, n), _unit(rng, n), _unit(rng, n)
    w = rng.exponential(size=n)
    w[rng.random(n) < 0.1] = 0.0  # zero-weight records, as training pools hold
    return make_records(pos, nrm, wo, wi, w)


def test_no_alloc_train_matches_allocating_step(torch_gpu):
    torch = torch_gpu
    from miniraytrace_amd import _lib as L
    from miniraytrace_amd.guide import GuideNetwork

    with GuideNetwork(max_batch=64, seed=5) as a, GuideNetwork(max_batch=64, seed=5) as b:
        assert np.array_equal(a.params, b.params)
        side = torch.cuda.Stream()  # a caller stream (default priority) for the first call
        stream = side.cuda_stream
        for step, (count, seed, use_stream) in enumerate(((6000, 11, True), (2500, 12, False)), 1):
            print(f"case: no-alloc training step {step}, {count} records", flush=True)
            rec = _records(count, seed)
            la = a.train_step(rec, lr=2e-3)
            d = torch.from_numpy(rec).cuda()
            torch.cuda.synchronize()
            lb = b.train_step_no_alloc(d.data_ptr(), count, lr=2e-3, stream=stream if use_stream else 0)
            torch.cuda.synchronize()
            # the loss is a float atomic sum over waves (order-dependent last
            # bits, for either entry point); the gradient is a fixed-order sum
            assert abs(la - lb) <= 1e-6 * abs(la), (la, lb)
            pa, pb = a.params, b.params
            assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32)), \
                f"{int((pa != pb).sum())} parameters differ after step {step}"
            ma, va, ta = a.adam_state()
            mb, vb, tb = b.adam_state()
            assert ta == tb == step
            assert np.array_equal(ma.view(np.uint32), mb.view(np.uint32))
            assert np.array_equal(va.view(np.uint32), vb.view(np.uint32))
            del d

        # beyond the fixed workspace: rejected before any device work
        n = L.MRTX_GUIDE_NOALLOC_MAX_RECORDS + 1
        big = torch.zeros((n, 16), dtype=torch.float32, device="cuda")
        with pytest.raises(L.MrtError) as err:
            b.train_step_no_alloc(big.data_ptr(), n)
        assert err.value.code == L.MRT_ERROR_INVALID_ARGUMENT
        assert "exceed" in L.last_error()
        assert b.adam_state()[2] == 2
