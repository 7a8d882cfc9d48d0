"""The map agreement metrics the parity tests gate on.  (CPU only.)"""
import numpy as np

from eman2_amd.metrics import fsc, max_rel_err


def _shell_filter(vol, keep):
    n = vol.shape[0]
    F = np.fft.rfftn(vol)
    k = np.sqrt(sum(np.meshgrid(np.fft.fftfreq(n) * n, np.fft.fftfreq(n) * n, np.fft.rfftfreq(n) * n,
                                indexing="ij")[i] ** 2 for i in range(3)))
    F[~keep(np.rint(k).astype(int))] = 0
    return np.fft.irfftn(F, s=vol.shape, axes=(0, 1, 2))


def test_fsc_identical_maps_is_one():
    a = np.random.default_rng(0).standard_normal((16, 16, 16))
    np.testing.assert_allclose(fsc(a, a), 1.0, atol=1e-12)


def test_fsc_shell_missing_in_one_map_fails():
    a = np.random.default_rng(1).standard_normal((16, 16, 16))
    b = _shell_filter(a, lambda s: s != 5)          # b lost shell 5 entirely
    f = fsc(b, a)
    assert abs(f[5]) < 0.2 and abs(fsc(a, b)[5]) < 0.2      # rounding dust left in b's shell 5
    assert np.all(np.delete(f, 5) > 0.999999)
    assert f.min() < 0.9999                          # the gate the parity tests use fails


def test_fsc_map_against_zero_map_is_zero():
    """every shell of a zero map is empty: 0/0 against a non-empty shell is a
    failure (0), not agreement"""
    a = np.random.default_rng(5).standard_normal((16, 16, 16))
    z = np.zeros_like(a)
    assert np.all(fsc(a, z) == 0.0) and np.all(fsc(z, a) == 0.0)
    assert np.all(fsc(z, z) == 1.0)


def test_fsc_shell_empty_in_both_maps_agrees():
    a = _shell_filter(np.random.default_rng(2).standard_normal((16, 16, 16)), lambda s: s != 3)
    assert fsc(a, a)[3] == 1.0


def test_fsc_sign_flip_and_noise():
    a = np.random.default_rng(3).standard_normal((16, 16, 16))
    assert np.all(fsc(a, -a)[1:] < -0.999999)
    b = a + 3 * np.random.default_rng(4).standard_normal(a.shape)
    assert fsc(a, b)[1:].mean() < 0.5


def test_max_rel_err():
    b = np.array([0.0, 2.0, -4.0])
    assert abs(max_rel_err(b + np.array([0.0, 0.0, 0.4]), b) - 0.1) < 1e-12
