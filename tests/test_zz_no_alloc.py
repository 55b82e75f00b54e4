"""The allocation-free entry point dqmc_apply_bop_noalloc against dqmc_apply_bop, through the C ABI.

The caller's matrices and fields live in device memory it owns (torch tensors here); the entry
point queues the product on the caller's stream with its intermediates in the library's static
device scratch.  Two calls in a row, the second on a smaller input, must each equal the
context-based dqmc_apply_bop bit for bit (the same kernels on the same factors), for all five
B-operations; an input beyond the scratch is refused with DQMC_EINVAL before any device work.
"""
import sys

import numpy as np
import pytest

import dqmc_amd as gpu

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def one_call(torch, nl, nt, beta, u, mu, nch, t, seed):
    n = nl * nl
    e = gpu.Engine(nl, nt, beta, u, mu=mu, pace=nt, n_chains=nch, delay_k=1)
    for c in range(nch):
        e.set_seed(c, seed + c)
        e.fields_random(c)
    e.init()
    fields = np.ascontiguousarray(np.stack([e.fields(c)[t] for c in range(nch)]))   # [chain][N] int8
    d_fields = torch.from_numpy(fields).to("cuda")
    rng = np.random.default_rng(seed)
    mats = rng.standard_normal((nch, n, n))
    stream = torch.cuda.current_stream()
    for op in range(5):
        want = [e.apply_bop(c, op, t, mats[c]) for c in range(nch)]
        d_x = torch.from_numpy(np.ascontiguousarray(np.stack([m.T for m in mats]))).to("cuda")   # column-major
        torch.cuda.synchronize()
        gpu.apply_bop_noalloc(nl, nt, beta, u, mu, 1.0, nch, op, d_fields.data_ptr(), d_x.data_ptr(),
                              stream.cuda_stream)
        torch.cuda.synchronize()
        got = d_x.cpu().numpy()
        for c in range(nch):
            assert np.array_equal(got[c].T, want[c]), "nl %d op %d chain %d: max diff %g" % (
                nl, op, c, np.abs(got[c].T - want[c]).max())
    e.close()


def test_noalloc_bops_equal_context_bops_two_calls():
    torch = pytest.importorskip("torch")
    say("case: noalloc nl=16 4 chains, then nl=6 3 chains")
    one_call(torch, 16, 10, 1.0, 4.0, -0.2, 4, 3, 700)
    one_call(torch, 6, 12, 1.2, 4.0, 0.3, 3, 11, 900)


def test_noalloc_refuses_input_beyond_scratch():
    say("case: noalloc oversize input")
    # 64 chains of 256 x 256 doubles are exactly the scratch, with no room for the constants
    with pytest.raises(gpu.DqmcError) as ei:
        gpu.apply_bop_noalloc(16, 10, 1.0, 4.0, 0.0, 1.0, 64, 0, 16, 16)
    assert ei.value.code == -1 and "exceed" in str(ei.value)
