"""Writes tests/golden/loader_chain.npz: outputs of the reference's own
loader-chain classes (RandomU, Hammerslay, Inverfc, BesselJ, QuietLoading,
ShotNoise, GaussHermite) compiled from the reference sources by
oracle/build_ref.sh into oracle/_ref/libg4ref.so.  Inputs and outputs only.

Usage: oracle/build_ref.sh && python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle_py as O  # noqa: E402

R = O.ref()
if R is None:
    sys.exit("oracle/_ref/libg4ref.so is missing; run oracle/build_ref.sh first")
p = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
out = {}
seeds = np.array([0, 1, 2, 7, 12345, 123456789], np.int64)
out["ran2_seeds"] = seeds
out["ran2"] = np.zeros((len(seeds), 200))
for i, s in enumerate(seeds):
    R.ref_ran2(int(s), 200, p(out["ran2"][i]))
starts = np.array([0, 987654321, 4000000000], np.int64)
out["halton_starts"] = starts
out["halton"] = np.zeros((6, len(starts), 100))
for b in range(6):
    for j, st in enumerate(starts):
        R.ref_hammersley(b, int(st), 100, p(out["halton"][b, j]))
ys = np.concatenate([np.linspace(1e-9, 2 - 1e-9, 401), [1e-300, 1e-100, 0.5, 1.0, 1.5]])
out["inverfc_y"] = ys
out["inverfc"] = np.array([R.ref_inverfc(float(y)) for y in ys])
xs = np.linspace(-25.0, 25.0, 251)
out["besselj_x"] = xs
out["besselj"] = np.array([[R.ref_besselj(n, float(x)) for x in xs] for n in range(5)])
sp = O.SliceParams(gamma=11357.8, delgam=1.13578, current=3000, ex=0.4e-6, ey=0.4e-6,
                   betax=15, betay=15, alphax=0.3, alphay=-0.2, xcen=1e-6, pxcen=0.1)
spv = np.array([getattr(sp, f[0]) for f in O.SliceParams._fields_])
out["quiet_params"] = spv
for islice in (0, 3, 1023):
    arrs = [np.zeros(256) for _ in range(6)]
    R.ref_quiet_load(*[p(a) for a in arrs], p(spv), 256, 4, 2 * np.pi, islice)
    out[f"quiet_{islice}"] = np.array(arrs)
h = R.ref_shotnoise_new(7, 0)
th0 = np.linspace(0.0, 2 * np.pi, 256, endpoint=False)
out["shot_theta_in"] = th0
for j in range(3):
    th = th0.copy()
    R.ref_shotnoise(h, p(th), 256, 4, 14126.0)
    out[f"shot_{j}"] = th
R.ref_shotnoise_free(h)
f = np.zeros(15 * 15 * 2)
R.ref_load_gauss(p(f), 15, 0.5e-3, 2.2616e-10, 5e3, 30e-6, 0.3, 0.2, 0.0, 0.0)
out["gauss15"] = f
np.savez_compressed(os.path.join(HERE, "loader_chain.npz"), **out)
print("wrote", os.path.join(HERE, "loader_chain.npz"))
