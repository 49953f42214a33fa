"""Write synth_golden.npz: steady states of the 64-species synthetic network
(prism3d_amd.network_synth) on six seeded cells, from scipy's BDF on the
oracle's restatement of the upstream RHS (oracle/generic_network.py) at
rtol 1e-10, atol 1e-24, with the analytic Jacobian.  With --wide, the same
for the 103-species network into synth_wide_golden.npz.  No upstream code
runs; this checks the GPU solve on networks the reference does not ship.

    python tests/golden/make_synth_golden.py [--wide] [--limit SECONDS]

The 103-species fixture was written with --wide --limit 300: one of its six
cells (~1e5 BDF steps on the GPU as well) did not finish in scipy within the
limit, and is recorded under 'dropped'.
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


def main(wide=False, limit=None):
    sp, rx = synthetic_network_wide() if wide else synthetic_network()
    net = GenericNetwork(sp, rx)
    y0 = initial(sp)
    t_end = 1e7 * 3.15576e7
    rng = np.random.RandomState(4)
    par = np.array([(10 ** rng.uniform(2, 5), 10 ** rng.uniform(1.2, 2.5),
                     10 ** rng.uniform(0, 3.5), rng.uniform(0, 6),
                     10 ** rng.uniform(-16.5, -15.5), 10 ** rng.uniform(-4, 0),
                     10 ** rng.uniform(-4, 0)) for _ in range(6)])
    # one process per cell, each single-threaded (the matrices are small);
    # with a time limit, a cell scipy has not finished by then is left out
    # and its parameters are kept in 'dropped'
    from multiprocessing import Pool
    with Pool(len(par)) as pool:
        jobs = [pool.apply_async(_cell, (wide, y0, p, t_end)) for p in par]
        out, keep = [], []
        for c, j in enumerate(jobs):
            try:
                out.append(j.get(timeout=limit))
                keep.append(c)
            except Exception as e:  # multiprocessing.TimeoutError
                print(f"cell {c} left out: {type(e).__name__}", flush=True)
        pool.terminate()
    dropped = par[[c for c in range(len(par)) if c not in keep]]
    par = par[keep]
    name = "synth_wide_golden.npz" if wide else "synth_golden.npz"
    np.savez_compressed(os.path.join(HERE, name), params=par, y0=y0,
                        t_end=t_end, y_end=np.array(out), rtol=1e-10, atol=1e-24,
                        dropped=dropped.reshape(-1, 7))


if __name__ == "__main__":
    args = sys.argv[1:]
    main(wide="--wide" in args,
         limit=float(args[args.index("--limit") + 1]) if "--limit" in args else None)
