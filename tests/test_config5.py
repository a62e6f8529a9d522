"""The uneven-box generator (host only)."""
import numpy as np

from piny_amd import water as w


def test_config5_boxes():
    cfgs = w.config5(nbox=16, max_mol=512)
    assert len(cfgs) == 16
    for c in cfgs:
        s = c.system
        assert 64 <= s.nmol <= 512
        assert s.pi_beads in (1, 4, 8, 16, 32, 64)
        assert s.settings.cutoff <= 0.5 * s.box - 0.05 + 1e-12
        assert c.x.size == s.natm * s.pi_beads
    # seeded
    again = w.config5(nbox=16, max_mol=512)
    assert all(np.array_equal(a.x, b.x) for a, b in zip(cfgs, again))
