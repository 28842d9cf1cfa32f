"""Host-side logic with no device work: the live-memory model of the
checkpointed environments (ttgpu.workloads.env_live_model)."""
import numpy as np

from ttgpu.workloads import env_live_model


def test_env_model_stride_one_is_one_full_set():
    N = 20
    env = np.full(2 * (N - 1), 1000.0)            # one sweep, equal sizes
    peak, _ = env_live_model(env, N, 1)
    # N environments (TL[0..site+1] and TR[site+1..N-1]) plus the one produced
    assert 8 * 1000 * N <= peak <= 8 * 1000 * (N + 2)


def test_env_model_stride_shrinks_the_peak_and_scales():
    N = 40
    rng = np.random.default_rng(0)
    env = rng.uniform(1, 5, 4 * (N - 1))          # two sweeps
    p1, _ = env_live_model(env, N, 1)
    p4, _ = env_live_model(env, N, 4)
    p4s, _ = env_live_model(env, N, 4, scale=4.0)
    assert p4 < 0.5 * p1
    assert np.isclose(p4s, 4 * p4)
    # every environment is counted at least once with stride 1
    assert p1 >= 8 * np.sort(env[-(N - 1):])[: N // 2].sum()
