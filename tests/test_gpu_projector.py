"""GPU parity of the gauss_fft projector (GaussFFTProjector::project3d,
libEM/projector.cpp:69-295) against the oracle: modes 0, 1 and 2, the
Fourier slice with the translation's phase ramp and the image output,
batches of transforms, a scaled transform.  Bar: max|a-b| <= 1e-4 max|b|
(the map's FFT is float on the GPU, double in the oracle)."""
import numpy as np
import pytest

from eman2_amd import synthetic as S
from eman2_amd.metrics import max_rel_err
from eman2_amd.transform import euler_eman_matrix

pytestmark = pytest.mark.gpu


def _mats(n, seed, shift=True, scale=None):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        m = euler_eman_matrix(*(rng.random(3) * (360, 180, 360))).astype(np.float32)
        if scale is not None:
            m[:, :3] *= np.float32(scale)
        if shift and i % 2:
            m[:2, 3] = rng.uniform(-4.0, 4.0, 2)
        out.append(m)
    return np.stack(out)


@pytest.mark.parametrize("P,mode,returnfft", [(48, 0, 1), (48, 1, 1), (48, 2, 1), (48, 0, 0), (96, 0, 1), (64, 2, 0)])
def test_gauss_fft_matches_oracle(oracle, P, mode, returnfft):
    from eman2_amd.projector import Projectors
    print("gauss_fft P=%d mode=%d returnfft=%d" % (P, mode, returnfft), flush=True)
    ph = S.phantom(20, P * 0.3, seed=P)
    vol = (S.phantom_volume(ph, P) + 0.01 * np.random.default_rng(1).standard_normal((P, P, P))).astype(np.float32)
    mats = _mats(6, P + mode)
    pr = Projectors.get("gauss_fft", {"mode": mode, "returnfft": returnfft})
    got = pr.project_batch(vol, mats)
    for i in range(len(mats)):
        want = oracle.gauss_fft_project(vol, mats[i], mode, returnfft)
        e = max_rel_err(got[i], want)
        assert e <= 1e-4, (i, e)
    one = Projectors.get("gauss_fft", {"transform": mats[1], "mode": mode, "returnfft": returnfft}).project3d(vol)
    assert max_rel_err(one.numpy(), got[1]) <= 1e-6


def test_gauss_fft_scaled_transform(oracle):
    from eman2_amd.projector import Projectors
    P = 48
    vol = np.random.default_rng(3).standard_normal((P, P, P)).astype(np.float32)
    mats = _mats(3, 7, shift=False, scale=1.2)
    got = Projectors.get("gauss_fft", {"returnfft": 1}).project_batch(vol, mats)
    for i in range(3):
        assert max_rel_err(got[i], oracle.gauss_fft_project(vol, mats[i], 0, 1)) <= 1e-4
