"""The device generator of the growth loop's divisions (one wavefront per
replicate, the 624-word twist split over the lanes) against the host
restatement of MT19937 + uniform_01 (itself pinned to numpy's MT19937 with the
reference's seeding and to the C++ standard's known answer in
test_sim_host.py): the uniform draws are bit-identical across several twists,
and the Box-Muller normals built from the next draws agree to the last bits of
the device's log and cos."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [0, 7, 42, 255, 2**32 - 1])
def test_device_stream_equals_host_stream(seed):
    from bfm import sim
    n = 2000   # three twists and a part
    print(f"device rng seed={seed} n={n}", flush=True)
    u_dev, z_dev = sim.rng_draws_device(seed, n)
    u_all = sim.rng_uniform(seed, 3 * n)
    assert np.array_equal(u_dev, u_all[:n])
    u, v = u_all[n::2], u_all[n + 1::2]
    z_ref = np.sqrt(-2 * np.log(u)) * np.cos(2 * np.pi * v)
    assert np.all(np.abs(z_dev - z_ref) <= 4e-16 * np.maximum(1.0, np.abs(z_ref)))
