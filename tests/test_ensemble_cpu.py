"""The seeded ensemble generator (dgsem_amd/ensemble.py), on the CPU."""
import numpy as np

from dgsem_amd import ensemble


def test_splitmix64_known_values():
    # SplitMix64 with seed 0: the first outputs of the published generator
    r = ensemble.SplitMix64(0)
    assert [r.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                                0x06C45D188009454F]


def test_ensemble_is_deterministic_and_in_range():
    a = ensemble.khi_ensemble(seed=2024, ncases=256)
    b = ensemble.khi_ensemble(seed=2024, ncases=256)
    assert len(a) == 256
    for x, y in zip(a, b):
        assert x.cells == y.cells and x.polydeg == y.polydeg
        assert np.array_equal(x.modes, y.modes)
    assert {c.polydeg for c in a} == {2, 3, 4, 5}
    assert {c.cells[0] for c in a} == {8, 16, 32, 64} == {c.cells[1] for c in a}
    m = np.stack([c.modes for c in a])
    assert m[..., :2].min() >= 1 and m[..., :2].max() <= 4
    assert 0 <= m[..., 2].min() and m[..., 2].max() < 2 * np.pi
    assert 0 <= m[..., 3].min() and m[..., 3].max() <= 0.02
    assert ensemble.khi_ensemble(seed=2025, ncases=4)[0].modes[0, 2] != a[0].modes[0, 2]


def test_khi_state_is_periodic_and_matches_unperturbed_formula():
    rng = np.random.default_rng(0)
    pts = rng.uniform(-1, 1, size=(50, 1, 2))
    modes = ensemble.khi_ensemble(seed=3, ncases=1)[0].modes
    def perturbation(x):
        # velocity perturbation alone: the base flow's shear layers are
        # periodic only up to tanh tails
        a, b = ensemble.khi_state(x, modes), ensemble.khi_state(x)
        return a[..., 1:3] / a[..., :1] - b[..., 1:3] / b[..., :1]

    for shift in ([2.0, 0.0], [0.0, 2.0]):
        assert np.allclose(perturbation(pts + np.array(shift)), perturbation(pts), atol=1e-13)
    base = ensemble.khi_state(pts)
    y = pts[..., 1]
    B = np.tanh(15 * y + 7.5) - np.tanh(15 * y - 7.5)
    assert np.allclose(base[..., 0], 0.5 + 0.75 * B)
    rho = base[..., 0]
    assert np.allclose(base[..., 1] / rho, 0.5 * (B - 1))
    p = 0.4 * (base[..., 3] - 0.5 * (base[..., 1] ** 2 + base[..., 2] ** 2) / rho)
    assert np.allclose(p, 1.0)
