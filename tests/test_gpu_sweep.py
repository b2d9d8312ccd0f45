"""The 256-point capillary sweep (energy U[50, 200] μJ × pressure U[0.5, 3]
bar from seed 1, half on the Nto = 8192 and half on the Nto = 16384 grid) in
one batch, checked point by point against the oracle: eight points spread
over the sweep and the slowest one (most step attempts) must reach the same
exit spectrum (relative spectral-energy error ≤ 1e-6 over sidx) with the same
number of step attempts.  This is Scans.runscan's per-point loop
(src/Scans.jl:275-289) run as one launch per grid class.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_seed1_sweep_against_oracle(pkg, orc, capi):
    print("case: 256-point sweep, seed 1", flush=True)
    from luna_jl_amd import configs, runtime
    insts = configs.scan_instances(256, seed=1)
    ctxs = [runtime.Context(c.holder) for c in insts]
    opts = capi.rk_opts(1.0, save_n=201)
    batch = runtime.ResidentBatch(ctxs, [c.Ew0 for c in insts])
    stats = batch.solve(opts)
    assert all(st.status == 0 for st in stats)
    slow = int(np.argmax([st.steps for st in stats]))
    picks = sorted(set(list(range(0, 256, 32)) + [slow]))
    assert len(picks) >= 9
    print(f"slowest point {slow}: {stats[slow].steps} attempts", flush=True)
    for i in picks:
        st_r, Ew_r, _, _, sr = orc.solve(insts[i].holder, insts[i].Ew0, opts, want_saves=False)
        assert st_r == 0
        sid = insts[i].sidx
        pa, pb = np.abs(batch.state(i)[sid]) ** 2, np.abs(Ew_r[sid]) ** 2
        err = np.linalg.norm(pa - pb) / np.linalg.norm(pb)
        print(f"point {i}: n_to {insts[i].params['n_to']}, steps {stats[i].steps}/{sr.steps}, err {err:.2e}",
              flush=True)
        assert stats[i].steps == sr.steps, i
        assert err <= 1e-6, (i, err)
