"""Phase timings of the driver (upstream self.timing) and the runtime
physics checks (prism3d/evaluate.py:176-223) on a solved grid."""

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def test_phase_timing_and_physics_checks(gpu):
    from prism3d_amd.solver3d import PDRSolverGPU
    from prism3d_amd.validation import physics_validation
    print("case: 8^3 fixture, 2 dust steps, timing + physics checks", flush=True)
    r = golden("run_golden.npz")
    sv = PDRSolverGPU(r['density'].reshape(8, 8, 8), float(r['box']), G0_external=100.0,
                      zeta_CR_0=2e-16, nside_rt=1)
    sv.run(max_iterations=4, convergence_tol=1e-9, dust_steps=2)
    for k in ('RT', 'shielding', 'dust', 'chemistry', 'thermal', 'dust_evolve'):
        assert sv.timing[k] > 0.0, k
    chk = physics_validation(sv)
    xh, xh2 = sv.x_HI.cpu().numpy(), sv.x_H2.cpu().numpy()
    xcp, xc, xco = sv.x_Cp.cpu().numpy(), sv.x_C.cpu().numpy(), sv.x_CO.cpu().numpy()
    t, fn = sv.T_gas.cpu().numpy(), sv.f_nano.cpu().numpy()
    h = xh + 2 * xh2
    assert chk['H_conservation']['mean'] == pytest.approx(np.mean(h), rel=1e-12)
    assert chk['H_conservation']['std'] == pytest.approx(np.std(h), rel=1e-9, abs=1e-15)
    assert chk['C_conservation']['ratio'] == pytest.approx(np.mean(xcp + xc + xco) / sv.x_c_total,
                                                           rel=1e-12)
    assert chk['dust_evolved']['f_nano_range'] == pytest.approx(fn.max() - fn.min(), abs=1e-15)
    corr = np.corrcoef(xcp.ravel(), xco.ravel())[0, 1]
    assert chk['chemical_stratification']['Cp_CO_correlation'] == pytest.approx(corr, abs=1e-9)
    assert chk['T_physical']['pass'] == bool((t > 0).all() and t.max() < 1e6)
    for k in ('H_conservation', 'C_conservation', 'T_physical'):
        assert chk[k]['pass'], (k, chk[k])
    print(chk['summary'], {k: v['pass'] for k, v in chk.items() if k != 'summary'}, flush=True)
