"""Writes tests/golden/faddeeva_golden.npz: w(z) = exp(-z^2) erfc(-iz) at
seeded points, from mpmath 1.3.0 (30 digits) and scipy.special.wofz.  These
are independent library values; neither the reference nor this engine is run.

Regions: the line-core domain |x| <= 130, y in [1e-7, 1e3] (log-uniform), the
real-axis-adjacent strip y < 1e-3, the imaginary axis (erfcx), and large |z|.
"""
import os

import mpmath as mp
import numpy as np
from scipy.special import wofz

mp.mp.dps = 30
rng = np.random.default_rng(20261203)
pts = []
pts += [(x, y) for x, y in zip(rng.uniform(-130, 130, 600), 10 ** rng.uniform(-7, 3, 600))]
pts += [(x, y) for x, y in zip(rng.uniform(-8, 8, 600), 10 ** rng.uniform(-7, np.log10(8), 600))]
pts += [(x, y) for x, y in zip(rng.uniform(0, 12, 300), 10 ** rng.uniform(-7, -3, 300))]
pts += [(0.0, y) for y in 10 ** rng.uniform(-7, 3, 100)]
pts += [(x, 0.0) for x in rng.uniform(-10, 10, 50)]
x = np.array([p[0] for p in pts])
y = np.array([p[1] for p in pts])
ref = np.array([complex(mp.exp(-mp.mpc(a, b) ** 2) * mp.erfc(-1j * mp.mpc(a, b))) for a, b in pts])
sp = wofz(x + 1j * y)
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "faddeeva_golden.npz")
np.savez_compressed(out, x=x, y=y, re=ref.real, im=ref.imag, scipy_re=sp.real, scipy_im=sp.imag)
print("wrote", out, len(x), "points")
