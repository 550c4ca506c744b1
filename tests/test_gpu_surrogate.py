"""The reflectance surrogate end to end on a small seeded set: batched FDTD
spectra of morpho-scale geometries train a 2-network ensemble in the
library's BF16 kernels; on held-out geometries it must beat the method
without a network (the nearest training geometry's spectrum), and the
fallback must hand exactly the flagged queries to the FDTD."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_surrogate_gain_and_fallback():
    from mpifdtd_amd import surrogate as S
    spec_tr = S.morpho_spec(seed=11, n=512)
    spec_te = S.morpho_spec(seed=12, n=48)
    Xtr, Ytr = S.generate(spec_tr)
    Xte, Yte = S.generate(spec_te)
    assert np.all(np.isfinite(Ytr)) and Ytr.min() >= 0 and Ytr.max() > 0
    scale = float(Ytr.mean())
    ens = S.Ensemble([S.N_FEATURES, 256, 256, S.N_ANGLES], k=2, batch=512, seed=3)
    hist = ens.fit(Xtr, (Ytr / scale).astype(np.float32), epochs=300, lr=1e-3)
    assert hist[-1] < hist[0] / 5
    mean, _ = ens.predict(Xte)
    err = S.rel_err(mean * scale, Yte)
    err_nn = S.rel_err(S.nearest_neighbour(Xtr, Ytr, Xte), Yte)
    print(f"held-out rel err: surrogate mean {err.mean():.4f} median {np.median(err):.4f}; "
          f"nearest neighbour mean {err_nn.mean():.4f} median {np.median(err_nn):.4f}")
    assert err.mean() < err_nn.mean()
    # threshold 0: every query goes to the FDTD, which reproduces the reference spectra
    out, use = S.predict_with_fallback(ens, spec_te[:8], scale, threshold=0.0)
    assert use.all() and np.array_equal(out, Yte[:8])
    # an unreachable threshold: none does, and the answer is the ensemble mean
    out, use = S.predict_with_fallback(ens, spec_te[:8], scale, threshold=1e30)
    assert not use.any()
    assert np.allclose(out, mean[:8] * scale, rtol=1e-6)
