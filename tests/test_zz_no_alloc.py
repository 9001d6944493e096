"""pp_overlap_gemm_noalloc (no device allocation; split-K partials in the
library's fixed scratch array) against pp_overlap_gemm through the C-ABI,
on two calls in a row, the second with a smaller input, in FP64 and FP32.
In FP64 both split K the same way within the scratch array, so the results
are identical, and match an FP64 numpy sum to 1e-12 normwise; in FP32 the
allocation-free call folds fewer, longer chunks (its FP32 partial sums
would not fit), so the two agree to 1e-6 normwise, the FP32 bound.  A call
beyond the scratch limit is refused with PP_ERR_ARG."""

import ctypes as ct

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pawpyseed_amd.core.batched import Context
    c = Context(0)
    yield c
    c.close()


def _inputs(seed, nkt, nbR, nbS, W, ld):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    CR = torch.randn((nkt, nbR, ld), dtype=torch.complex64, device="cuda", generator=g)
    CS = torch.randn((nkt, nbS, ld), dtype=torch.complex64, device="cuda", generator=g)
    nw = np.array([W - 37 * k for k in range(nkt)], np.int32)
    return CR, CS, nw


def _call(fn, ctx, nwk, nspin, flip, nw, CR, CS, out, prec):
    from pawpyseed_amd.core._lib import c_int_p, lib, ptr
    nkt, nbR, ld = CR.shape
    nbS = CS.shape[1]
    return getattr(lib(), fn)(ctx.h, nwk, nspin, flip, ptr(nw, c_int_p), ct.c_void_p(CR.data_ptr()), ld, nbR,
                              ct.c_void_p(CS.data_ptr()), ld, nbS, ct.c_void_p(out.data_ptr()), prec)


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_overlap_gemm(ctx, prec):
    from pawpyseed_amd.core._lib import lib
    assert lib().pp_noalloc_scratch_bytes() == 64 << 20
    for call, (nwk, nspin, flip, nbR, nbS, W, seed) in enumerate(((2, 2, 1, 512, 512, 20000, 1),
                                                                   (1, 2, 1, 96, 200, 7000, 2))):
        ld = -(-W // 64) * 64
        CR, CS, nw = _inputs(seed, nwk * nspin, nbR, nbS, W, ld)
        ref = torch.zeros((nwk * nspin, nbR, nbS), dtype=torch.complex128, device="cuda")
        got = torch.full_like(ref, complex(7.0, -7.0))
        ctx.synchronize()
        assert _call("pp_overlap_gemm", ctx, nwk, nspin, flip, nw, CR, CS, ref, prec) == 0
        assert _call("pp_overlap_gemm_noalloc", ctx, nwk, nspin, flip, nw, CR, CS, got, prec) == 0
        ctx.synchronize()
        diff = ((got - ref).abs().max() / ref.abs().max()).item()
        print(f"call {call} prec {prec}: max |noalloc - overlap_gemm| / max |overlap_gemm| = {diff:.3e}", flush=True)
        if prec == 0:
            assert torch.equal(got, ref)  # the same FP64 split
        else:
            assert diff < 1e-6  # FP32: another split of the same sums
        if prec == 0:
            kR = [k + nwk if k < nwk else k - nwk for k in range(nwk * nspin)]
            for k in range(nwk * nspin):
                a = CR[kR[k], :, : nw[k]].cpu().numpy().astype(np.complex128)
                b = CS[k, :, : nw[k]].cpu().numpy().astype(np.complex128)
                want = a.conj() @ b.T
                err = np.abs(got[k].cpu().numpy() - want).max() / np.abs(want).max()
                assert err < 1e-12, (call, k, err)


def test_noalloc_refuses_beyond_scratch(ctx):
    from pawpyseed_amd.core._lib import PP_ERR_ARG, lib
    CR, CS, nw = _inputs(3, 1, 1500, 1500, 64, 64)
    out = torch.zeros((1, 1500, 1500), dtype=torch.complex128, device="cuda")
    assert _call("pp_overlap_gemm_noalloc", ctx, 1, 1, 0, nw, CR, CS, out, 0) == PP_ERR_ARG
    assert b"scratch limit" in lib().pp_last_error()
    assert _call("pp_overlap_gemm_noalloc", ctx, 1, 1, 0, nw, CR, CS, out, 2) == PP_ERR_ARG
