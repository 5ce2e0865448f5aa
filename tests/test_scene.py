"""Scene generators: the batched-scene workload is deterministic and within its stated ranges."""
import numpy as np

from fluid_engine import scene


def test_batch_scenes_deterministic_and_in_range():
    a = scene.batch_scenes(count=64, seed=0)
    b = scene.batch_scenes(count=64, seed=0)
    assert len(a) == 64
    for wa, wb in zip(a, b):
        assert np.array_equal(wa.positions, wb.positions)
        assert 3000 <= wa.n <= 300000
        p = wa.params
        assert 950.0 <= p.target_density <= 1050.0
        assert 0.005 <= p.viscosity <= 0.05 + 1e-7
        s = np.float32(0.02)
        assert abs(p.particle_mass - np.float32(p.target_density) * s * s * s) <= 1e-9
        edge = (wa.n / 32768.0) ** (1.0 / 3.0)
        assert abs(p.box_max[0] - edge) < 1e-5
        # the block starts at the box corner and fits in the box
        assert (wa.positions.min(axis=0) >= 0).all()
        assert (wa.positions.max(axis=0) <= np.array(p.box_max[:3])).all()
    n = np.array([w.n for w in a])
    assert n.min() < 8000 and n.max() > 150000     # log-uniform over [4096, 262144]
