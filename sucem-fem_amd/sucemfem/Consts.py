"""Physical constants -- values and evaluation order of sucemfem/Consts.py:24-33."""
from __future__ import division

import numpy as N

pi = N.pi
e = N.e

mu0 = N.pi * 4 * 1e-7                   # per definition
c0 = 299792458                          # per definition
eps0 = 10 ** 7 / 4 / N.pi / c0 / c0     # per definition
Z0 = mu0 * c0
