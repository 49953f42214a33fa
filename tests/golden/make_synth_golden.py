"""Write synth_golden.npz: steady states of the 64-species synthetic network
(prism3d_amd.network_synth) on six seeded cells, from scipy's BDF on the
oracle's restatement of the upstream RHS (oracle/generic_network.py) at
rtol 1e-10, atol 1e-24, with the analytic Jacobian.  With --wide, the same
for the 103-species network into synth_wide_golden.npz.  No upstream code
runs; this checks the GPU solve on networks the reference does not ship.

    python tests/golden/make_synth_golden.py [--wide]
"""
import os
import sys

for _v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ[_v] = "1"

import numpy as np  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.generic_network import GenericNetwork, integrate  # noqa: E402
from prism3d_amd.network_synth import synthetic_network, synthetic_network_wide  # noqa: E402

INIT = (('H', 1.0), ('C+', 1.4e-4), ('O', 3e-4), ('N', 7.5e-5), ('He', 0.1), ('S+', 1e-6),
        ('Si+', 1e-7), ('Fe+', 1e-8))


def initial(species):
    y0 = np.full(len(species), 1e-30)
    for s, v in INIT:
        y0[species.index(s)] = v
    y0[species.index('e-')] = 1.4e-4 + 1e-6 + 1e-7 + 1e-8
    return y0


def _cell(wide, y0, p, t_end):
    sp, rx = synthetic_network_wide() if wide else synthetic_network()
    y = integrate(GenericNetwork(sp, rx), y0, *p, t_end, rtol=1e-10, atol=1e-24)
    print("cell done", flush=True)
    return y


def main(wide=False):
    sp, rx = synthetic_network_wide() if wide else synthetic_network()
    net = GenericNetwork(sp, rx)
    y0 = initial(sp)
    t_end = 1e7 * 3.15576e7
    rng = np.random.RandomState(4)
    par = np.array([(10 ** rng.uniform(2, 5), 10 ** rng.uniform(1.2, 2.5),
                     10 ** rng.uniform(0, 3.5), rng.uniform(0, 6),
                     10 ** rng.uniform(-16.5, -15.5), 10 ** rng.uniform(-4, 0),
                     10 ** rng.uniform(-4, 0)) for _ in range(6)])
    # one process per cell, each single-threaded (the matrices are small)
    from multiprocessing import Pool
    with Pool(len(par)) as pool:
        out = pool.starmap(_cell, [(wide, y0, p, t_end) for p in par])
    name = "synth_wide_golden.npz" if wide else "synth_golden.npz"
    np.savez_compressed(os.path.join(HERE, name), params=par, y0=y0,
                        t_end=t_end, y_end=np.array(out), rtol=1e-10, atol=1e-24)


if __name__ == "__main__":
    main(wide="--wide" in sys.argv[1:])
