"""Many independent clouds advanced together.

Each cloud keeps its own ``PDRSolverGPU`` state and runs the grid work of an
inner iteration on its own grid (columns, shielding, dust, then damping and
the thermal balance).  The per-cell chemistry of every cloud still iterating
goes to ONE ``p3d_chem_solve`` launch over their concatenated cells, so the
persistent work-queue kernel sees the whole ensemble: cells of calm and
strongly irradiated clouds, whose step counts differ by orders of
magnitude, share the device queue instead of each small grid under-filling
the GPU on its own.

The chemistry is independent per cell, so every cloud follows the same
sequence of iterations as in its own ``run()`` (the upstream per-grid loop,
solver_3d.py:197-463): its inner loop stops at its own convergence, after
which its cells leave the shared launch.
"""

import sys

import torch

from .chemistry import ChemistrySolver


class CloudBatch:
    """Inner fixed-point iterations of several PDRSolverGPU clouds with one
    chemistry launch per iteration.  Every cloud must use
    chemistry='network' with the same chemistry settings (mode, horizon,
    tolerances); the kernel choice of the shared launch comes from the
    arguments here."""

    def __init__(self, clouds, chem_variant=0, chem_waves_per_cu=0, chem_rescue_frac=0.0):
        self.clouds = list(clouds)
        if not self.clouds:
            raise ValueError("a batch needs at least one cloud")
        c0 = self.clouds[0]
        key = lambda c: (c.chemistry, c.chem_mode, c.t_chem, c.chem_rtol, c.chem_atol,
                         tuple(c.species), c.fixed_T is None, c.accelerator is None)
        for c in self.clouds:
            if c.chemistry != 'network' or c.accelerator is not None:
                raise ValueError("batched clouds must use the network chemistry")
            if key(c) != key(c0) or c.dev != c0.dev:
                raise ValueError("batched clouds must share chemistry settings and device")
        self.cfg = (int(chem_variant), int(chem_waves_per_cu), float(chem_rescue_frac))
        self.n_species = c0.y.shape[0]
        self.n_cells = sum(c.n_cells for c in self.clouds)
        dev = c0.dev
        # flat storage: the first k cells of the batch view as contiguous
        # (n_species, k) and (7, k) arrays
        self._y = torch.empty(self.n_species * self.n_cells, dtype=torch.float64, device=dev)
        self._inp = torch.empty(7 * self.n_cells, dtype=torch.float64, device=dev)
        self.chem = ChemistrySolver()
        self.chem_stats = None
        self.launch_cells = []  # cells of every shared chemistry launch

    def chemistry_step(self, active):
        """One chemistry launch over the cells of the clouds in `active`."""
        k = sum(c.n_cells for c in active)
        y = self._y[:self.n_species * k].view(self.n_species, k)
        inp = self._inp[:7 * k].view(7, k)
        off = 0
        for c in active:
            n = c.n_cells
            if c.chem_mode == 'integrate':
                c.y.copy_(c.y_init)
            y[:, off:off + n].copy_(c.y)
            for r, a in enumerate((c.density, c.T_gas, c.G0, c.A_V, c.zeta_CR, c.fsh_H2,
                                   c.fsh_CO)):
                inp[r, off:off + n].copy_(a.reshape(-1))
            off += n
        _, stats = self.chem.solve_steady_batch(y, *[inp[r] for r in range(7)],
                                                opts=active[0].chem_opts(self.cfg), sync=False)
        off = 0
        for c in active:
            n = c.n_cells
            c.y.copy_(y[:, off:off + n])
            c.chem_stats = stats[off:off + n]
            off += n
        self.chem_stats = stats
        self.launch_cells.append(k)

    def run(self, max_iterations=20, convergence_tol=0.05, dust_steps=1, verbose=False):
        """The per-cloud run() loop (solver_3d.py:197-463) for all clouds at
        once; returns each cloud's converged flag for the last dust step."""
        n = len(self.clouds)
        converged = [False] * n
        for dust_step in range(int(dust_steps)):
            converged = [False] * n
            w_damp = []
            for c in self.clouds:
                if dust_step > 0:
                    c.evolve_dust()
                w_damp.append(torch.full(c.shape, 0.5, dtype=torch.float64, device=c.dev))
            for it in range(max_iterations):
                idx = [i for i in range(n) if not converged[i]]
                if not idx:
                    break
                active = [self.clouds[i] for i in idx]
                first = dust_step == 0 and it == 0
                prev = [c._before_chemistry(first) for c in active]
                self.chemistry_step(active)
                for j, (i, c) in enumerate(zip(idx, active)):
                    old_T, old_H2, old_CO, w_damp[i] = c._after_chemistry(it, w_damp[i],
                                                                          *prev[j])
                    rec = c._convergence(old_T, old_H2, old_CO, convergence_tol)
                    c._resolve_timing()
                    rec['dust_step'] = dust_step
                    rec['iteration'] = it
                    rec['f_nano_mean'] = c.f_nano_mean()
                    c.convergence_history.append(rec)
                    if verbose:
                        print(f"  cloud {i} {dust_step}.{it:2d} dT={rec['dT']:.2e} "
                              f"dH2={rec['dH2']:.2e} dCO={rec['dCO']:.2e}",
                              file=sys.stderr, flush=True)
                    if rec['delta_max'] < convergence_tol and it > 2:
                        converged[i] = True
        return converged
