"""Host threads driving distinct contexts at the same time.

The reference's Cython wrappers hold the GIL, so its C core never runs two
calls at once (pawpyc.pyx). This C-ABI promises thread safety across distinct
handles: each context owns one HIP stream and its workspaces, the error
state is per thread, and the compat layer's shared default context is
guarded by a mutex. ctypes releases the GIL around every library call, so
the threads below run their setups concurrently.
"""

import threading

import numpy as np
import pytest

from helpers import device_inputs, normwise, oracle_projections
from pawpyseed_amd import synthetic

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_two_threads_two_contexts():
    """Two threads, each with its own context, set up the projections of the
    64-site 60^3 cell (16 bands; FP64 in one thread, FP32 in the other)
    three times over while the other does the same. Every result equals the
    same call made alone, bit for bit (the kernels are deterministic), and
    the FP64 one matches the oracle to 1e-9."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pawpyseed_amd.core import batched as B

    wl = synthetic.config0()
    nb, reps = 16, 3
    Gs = wl.ensure_gvecs()
    Cs = wl.coefficients("S", nband=nb)
    Cd, Gd, gofs, _ = device_inputs(Gs, Cs)
    labels = np.zeros(len(wl.frac_S), np.int32)
    precs = ["fp64", "fp32"]
    ctxs = [B.Context(0) for _ in precs]
    plans = [B.SpherePlan(c, wl.lattice, wl.fftg, wl.frac_S, labels, [wl.species], p) for c, p in zip(ctxs, precs)]
    alone = []
    for c, pl, p in zip(ctxs, plans, precs):
        o = torch.zeros((len(Gs), nb, pl.ptot), dtype=torch.complex128, device="cuda")
        torch.cuda.synchronize()
        B.setup_projections(c, pl, wl.kpts, gofs, Gd, Cd, o, p)
        alone.append(o.cpu().numpy())
    outs = [[torch.zeros((len(Gs), nb, pl.ptot), dtype=torch.complex128, device="cuda") for _ in range(reps)]
            for pl in plans]
    # the contexts' streams do not wait for torch's default stream
    torch.cuda.synchronize()
    errors = []
    start = threading.Barrier(len(precs))

    def work(i):
        try:
            start.wait(timeout=60)
            for r in range(reps):
                B.setup_projections(ctxs[i], plans[i], wl.kpts, gofs, Gd, Cd, outs[i][r], precs[i])
        except Exception as e:  # reported by the main thread
            errors.append((precs[i], repr(e)))

    print("threads: fp64 and fp32 setups on two contexts", flush=True)
    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(precs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a thread did not finish"
    assert not errors, errors
    torch.cuda.synchronize()
    for i, p in enumerate(precs):
        for r in range(reps):
            assert np.array_equal(outs[i][r].cpu().numpy(), alone[i]), (p, r)
    want, _ = oracle_projections(wl, wl.frac_S, Cs)
    err = normwise(alone[0], want)
    print("threads: fp64 against the oracle", err)
    assert err < 1e-9
    for pl in plans:
        pl.close()
    for c in ctxs:
        c.close()
