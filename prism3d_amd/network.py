"""The bundled PDR chemical network and its upload to the device.

Species and reactions restate prism3d/chemistry/network.py:186-896 (75
reactions; rate parameters alpha, beta, gamma as published there) with the
31 tracked species of prism3d/chemistry/solver.py:28-36.  PAH, PAH+ and PAH-
are not tracked: PAH- enters two-body rates through its closed-form
abundance (solver.py:74-87), PAH is dropped as a product.

``Network`` holds the tables on the host; ``DeviceNetwork`` uploads them once
through p3d_network_create and owns the handle.
"""

import ctypes

import numpy as np

from . import _lib

TRACKED = ['e-', 'H', 'H2', 'H+', 'H-', 'H2+', 'H3+',
           'C+', 'C', 'CO', 'CH', 'CH+', 'CH2+',
           'O', 'O+', 'OH', 'OH+', 'H2O', 'H2O+', 'H3O+',
           'O2',
           'CO+', 'HCO+',
           'S', 'S+', 'Si', 'Si+', 'Fe', 'Fe+',
           'He', 'He+']

# upstream reaction-type names -> device rate forms (solver.py:104-133)
_TYPE = {'two_body': _lib.P3D_RX_TWO_BODY, 'recomb': _lib.P3D_RX_TWO_BODY,
         'charge_exchange': _lib.P3D_RX_TWO_BODY, 'grain_recomb': _lib.P3D_RX_TWO_BODY,
         'photodiss': _lib.P3D_RX_PHOTO, 'photoion': _lib.P3D_RX_PHOTO,
         'cosmic_ray': _lib.P3D_RX_CR, 'cr_photo': _lib.P3D_RX_CR,
         'grain_surface': _lib.P3D_RX_GRAIN}

# (name, type, reactants, products, alpha, beta, gamma) in upstream order
REACTIONS = [
    ('H2_formation_grain', 'grain_surface', ['H', 'H'], ['H2'], 3.0e-17, 0.5, 0.0),
    ('H2_photodiss', 'photodiss', ['H2'], ['H', 'H'], 3.3e-11, 0.0, 3.74),
    ('H2_CR_ionization', 'cosmic_ray', ['H2'], ['H2+', 'e-'], 2.5, 0.0, 0.0),
    ('H_CR_ionization', 'cosmic_ray', ['H'], ['H+', 'e-'], 0.46, 0.0, 0.0),
    ('H_recombination', 'recomb', ['H+', 'e-'], ['H'], 2.753e-14, -0.745, 0.0),
    ('H_minus_formation', 'two_body', ['H', 'e-'], ['H-'], 1.4e-18, 0.928, 0.0),
    ('H2_gasphase_Hminus', 'two_body', ['H-', 'H'], ['H2', 'e-'], 4.0e-9, -0.02, 0.0),
    ('C_photoionization', 'photoion', ['C'], ['C+', 'e-'], 3.0e-10, 0.0, 3.0),
    ('Cp_recombination', 'recomb', ['C+', 'e-'], ['C'], 4.67e-12, -0.6, 0.0),
    ('CO_photodiss', 'photodiss', ['CO'], ['C', 'O'], 2.0e-10, 0.0, 3.53),
    ('CO_form_CpOH', 'two_body', ['C+', 'OH'], ['CO', 'H+'], 7.7e-10, -0.5, 0.0),
    ('HCOp_form_CpH2O', 'two_body', ['C+', 'H2O'], ['HCO+', 'H'], 9.0e-10, 0.0, 0.0),
    ('CO_form_COH', 'two_body', ['C', 'OH'], ['CO', 'H'], 1.0e-10, 0.0, 0.0),
    ('CO_form_CHO', 'two_body', ['CH', 'O'], ['CO', 'H'], 6.59e-11, 0.0, 0.0),
    ('CHp_form', 'two_body', ['C+', 'H2'], ['CH+', 'H'], 1.0e-10, 0.0, 4640.0),
    ('CH2p_form', 'two_body', ['CH+', 'H2'], ['CH2+', 'H'], 1.2e-9, 0.0, 0.0),
    ('CHp_recomb', 'recomb', ['CH+', 'e-'], ['C', 'H'], 1.5e-7, -0.42, 0.0),
    ('O_Hp_CE', 'charge_exchange', ['O', 'H+'], ['O+', 'H'], 7.31e-10, 0.23, 225.9),
    ('Op_H_CE', 'charge_exchange', ['O+', 'H'], ['O', 'H+'], 5.66e-10, 0.36, -8.6),
    ('OHp_form_OH3p', 'two_body', ['O', 'H3+'], ['OH+', 'H2'], 8.4e-10, 0.0, 0.0),
    ('H2Op_form', 'two_body', ['OH+', 'H2'], ['H2O+', 'H'], 1.01e-9, 0.0, 0.0),
    ('H3Op_form', 'two_body', ['H2O+', 'H2'], ['H3O+', 'H'], 6.4e-10, 0.0, 0.0),
    ('H3Op_recomb_OH', 'recomb', ['H3O+', 'e-'], ['OH', 'H', 'H'], 1.08e-7, -0.5, 0.0),
    ('H3Op_recomb_H2O', 'recomb', ['H3O+', 'e-'], ['H2O', 'H'], 6.02e-8, -0.5, 0.0),
    ('OH_photodiss', 'photodiss', ['OH'], ['O', 'H'], 3.5e-10, 0.0, 1.7),
    ('H2O_photodiss', 'photodiss', ['H2O'], ['OH', 'H'], 5.9e-10, 0.0, 1.7),
    ('OH_form_OH2', 'two_body', ['O', 'H2'], ['OH', 'H'], 3.14e-13, 2.7, 3150.0),
    ('OH_destr_OHH', 'two_body', ['OH', 'H'], ['O', 'H2'], 6.99e-14, 2.8, 1950.0),
    ('H2O_form_OHH2', 'two_body', ['OH', 'H2'], ['H2O', 'H'], 2.05e-12, 1.52, 1736.0),
    ('HCOp_form_COH3p', 'two_body', ['CO', 'H3+'], ['HCO+', 'H2'], 1.36e-9, 0.0, 0.0),
    ('HCOp_recomb', 'recomb', ['HCO+', 'e-'], ['CO', 'H'], 2.4e-7, -0.69, 0.0),
    ('HCOp_photodiss', 'photodiss', ['HCO+'], ['CO+', 'H'], 1.5e-10, 0.0, 2.5),
    ('H3p_form', 'two_body', ['H2+', 'H2'], ['H3+', 'H'], 2.08e-9, 0.0, 0.0),
    ('H3p_recomb_H2', 'recomb', ['H3+', 'e-'], ['H2', 'H'], 4.36e-8, -0.52, 0.0),
    ('H3p_recomb_3H', 'recomb', ['H3+', 'e-'], ['H', 'H', 'H'], 4.36e-8, -0.52, 0.0),
    ('CHp_form_H3pC', 'two_body', ['H3+', 'C'], ['CH+', 'H2'], 2.0e-9, 0.0, 0.0),
    ('S_photoion', 'photoion', ['S'], ['S+', 'e-'], 6.1e-10, 0.0, 2.6),
    ('Sp_recomb', 'recomb', ['S+', 'e-'], ['S'], 3.9e-11, -0.63, 0.0),
    ('Si_photoion', 'photoion', ['Si'], ['Si+', 'e-'], 4.5e-9, 0.0, 2.6),
    ('Sip_recomb', 'recomb', ['Si+', 'e-'], ['Si'], 4.26e-12, -0.62, 0.0),
    ('Fe_photoion', 'photoion', ['Fe'], ['Fe+', 'e-'], 2.6e-9, 0.0, 2.8),
    ('Fep_recomb', 'recomb', ['Fe+', 'e-'], ['Fe'], 3.7e-12, -0.65, 0.0),
    ('COp_H', 'two_body', ['CO+', 'H'], ['CO', 'H+'], 7.5e-10, 0.0, 0.0),
    ('COp_H2_HCOp', 'two_body', ['CO+', 'H2'], ['HCO+', 'H'], 7.5e-10, 0.0, 0.0),
    ('He_CR_ion', 'cosmic_ray', ['He'], ['He+', 'e-'], 0.5, 0.0, 0.0),
    ('Hep_CO_destruct', 'two_body', ['He+', 'CO'], ['C+', 'O', 'He'], 1.6e-9, 0.0, 0.0),
    ('Hep_recomb', 'recomb', ['He+', 'e-'], ['He'], 1.0e-12, -0.64, 0.0),
    ('Cp_PAH_recomb', 'grain_recomb', ['C+', 'PAH-'], ['C', 'PAH'], 4.4e-12, 0.5, 0.0),
    ('Sp_PAH_recomb', 'grain_recomb', ['S+', 'PAH-'], ['S', 'PAH'], 4.4e-12, 0.5, 0.0),
    ('Hp_PAH_recomb', 'grain_recomb', ['H+', 'PAH-'], ['H', 'PAH'], 4.4e-12, 0.5, 0.0),
    ('Sip_PAH_recomb', 'grain_recomb', ['Si+', 'PAH-'], ['Si', 'PAH'], 4.4e-12, 0.5, 0.0),
    ('Fep_PAH_recomb', 'grain_recomb', ['Fe+', 'PAH-'], ['Fe', 'PAH'], 4.4e-12, 0.5, 0.0),
    ('CO_CR_photodiss', 'cr_photo', ['CO'], ['C', 'O'], 6.0, 0.0, 0.0),
    ('H2O_CR_photodiss', 'cr_photo', ['H2O'], ['OH', 'H'], 6.0, 0.0, 0.0),
    ('OH_CR_photodiss', 'cr_photo', ['OH'], ['O', 'H'], 5.0, 0.0, 0.0),
    ('C_CR_photoion', 'cr_photo', ['C'], ['C+', 'e-'], 3.9, 0.0, 0.0),
    ('O2_form_OOH', 'two_body', ['O', 'OH'], ['O2', 'H'], 3.5e-11, 0.0, 0.0),
    ('CO_form_CO2', 'two_body', ['C', 'O2'], ['CO', 'O'], 5.56e-11, 0.41, -26.9),
    ('O2_photodiss', 'photodiss', ['O2'], ['O', 'O'], 7.9e-10, 0.0, 1.8),
    ('O2_CR_photodiss', 'cr_photo', ['O2'], ['O', 'O'], 6.0, 0.0, 0.0),
    ('COp_form_O2Cp', 'two_body', ['O2', 'C+'], ['CO+', 'O'], 3.8e-10, 0.0, 0.0),
    ('CH_destr_H', 'two_body', ['CH', 'H'], ['C', 'H2'], 2.7e-11, 0.38, 0.0),
    ('CH_photodiss', 'photodiss', ['CH'], ['C', 'H'], 9.2e-10, 0.0, 1.7),
    ('CH2p_recomb_CH', 'recomb', ['CH2+', 'e-'], ['CH', 'H'], 1.6e-7, -0.6, 0.0),
    ('CH2p_recomb_C', 'recomb', ['CH2+', 'e-'], ['C', 'H', 'H'], 4.0e-7, -0.6, 0.0),
    ('HCOp_form_CHO', 'two_body', ['CH', 'O'], ['HCO+', 'e-'], 2.0e-11, 0.44, 0.0),
    ('COp_recomb', 'recomb', ['CO+', 'e-'], ['C', 'O'], 2.0e-7, -0.48, 0.0),
    ('CO_form_CH2O', 'two_body', ['C', 'H2O'], ['CO', 'H', 'H'], 1.2e-10, 0.0, 0.0),
    ('H2_coll_dissoc_H', 'two_body', ['H2', 'H'], ['H', 'H', 'H'], 6.67e-12, 0.0, 52000.0),
    ('H2_coll_dissoc_H2', 'two_body', ['H2', 'H2'], ['H2', 'H', 'H'], 5.5e-29, 0.0, 52000.0),
    ('Fep_grain_recomb', 'grain_recomb', ['Fe+', 'e-'], ['Fe'], 5.0e-12, -0.5, 0.0),
    ('Sip_grain_recomb', 'grain_recomb', ['Si+', 'e-'], ['Si'], 5.0e-12, -0.5, 0.0),
    ('OHp_recomb', 'recomb', ['OH+', 'e-'], ['O', 'H'], 3.75e-8, -0.5, 0.0),
    ('H2Op_recomb_O', 'recomb', ['H2O+', 'e-'], ['O', 'H', 'H'], 3.05e-7, -0.5, 0.0),
    ('H2Op_recomb_OH', 'recomb', ['H2O+', 'e-'], ['OH', 'H'], 8.6e-8, -0.5, 0.0),
]

# weights of the post-solve corrections (solver.py:227-388)
H_MINOR = {'H+': 1, 'H-': 1, 'H2+': 2, 'H3+': 3, 'OH': 1, 'OH+': 1, 'H2O': 2, 'H2O+': 2,
           'H3O+': 3, 'CH': 1, 'CH+': 1, 'HCO+': 1}
C_MINOR = {'CH': 1, 'CH+': 1, 'CH2+': 1, 'CO+': 1, 'HCO+': 1}
O_MOL = {'CO': 1, 'OH': 1, 'H2O': 1, 'O+': 1, 'OH+': 1, 'H2O+': 1, 'H3O+': 1, 'CO+': 1,
         'HCO+': 1, 'O2': 2}
CATIONS = ['H+', 'C+', 'S+', 'Si+', 'Fe+', 'O+', 'H2+', 'H3+', 'OH+', 'H2O+', 'H3O+', 'CH+',
           'CH2+', 'CO+', 'HCO+', 'He+']
ROLES = ['H', 'H2', 'H+', 'C+', 'C', 'CO', 'O']


class Network:
    """Host tables of a network over ``species`` (tracked, in state order)."""

    def __init__(self, species=None, reactions=None, shielded=('H2_photodiss', 'CO_photodiss')):
        self.species = list(species or TRACKED)
        self.reactions = list(reactions or REACTIONS)
        self.idx = {s: i for i, s in enumerate(self.species)}
        self.n_species = len(self.species)
        self.n_reactions = len(self.reactions)
        ns, nr = self.n_species, self.n_reactions
        self.rtype = np.zeros(nr, dtype=np.int32)
        self.alpha = np.zeros(nr)
        self.beta = np.zeros(nr)
        self.gamma = np.zeros(nr)
        self.reactants = np.full(2 * nr, _lib.P3D_SP_NONE, dtype=np.int32)
        self.products = np.full(3 * nr, _lib.P3D_SP_NONE, dtype=np.int32)
        self.names = []
        for r, (name, typ, reac, prod, a, b, g) in enumerate(self.reactions):
            t = _TYPE[typ]
            if t == _lib.P3D_RX_PHOTO and name == shielded[0]:
                t = _lib.P3D_RX_PHOTO_FH2
            elif t == _lib.P3D_RX_PHOTO and name == shielded[1]:
                t = _lib.P3D_RX_PHOTO_FCO
            self.rtype[r] = t
            self.alpha[r], self.beta[r], self.gamma[r] = a, b, g
            self.names.append(name)
            for m, s in enumerate(reac[:2]):
                self.reactants[2 * r + m] = self._code(s)
            for m, s in enumerate(prod[:3]):
                self.products[3 * r + m] = self._code(s)
        self.e_index = self.idx['e-']
        self.roles = np.array([self.idx.get(s, -1) for s in ROLES], dtype=np.int32)
        w = np.zeros((4, ns))
        for tab, row in ((H_MINOR, 0), (C_MINOR, 1), (O_MOL, 2)):
            for s, v in tab.items():
                if s in self.idx:
                    w[row, self.idx[s]] = v
        for s in CATIONS:
            if s in self.idx:
                w[3, self.idx[s]] = 1.0
        self.post_w = w

    def _code(self, s):
        if s in self.idx:
            return self.idx[s]
        if s == 'PAH-':
            return _lib.P3D_SP_PAH_MINUS
        return _lib.P3D_SP_NONE

    def stoichiometry(self):
        """(n_species, n_reactions) net stoichiometric matrix of tracked species."""
        S = np.zeros((self.n_species, self.n_reactions))
        for r, (_, _, reac, prod, *_rest) in enumerate(self.reactions):
            for s in reac:
                if s in self.idx:
                    S[self.idx[s], r] -= 1
            for s in prod:
                if s in self.idx:
                    S[self.idx[s], r] += 1
        return S


class DeviceNetwork:
    """A network uploaded to the GPU (p3d_network_create)."""

    def __init__(self, net=None):
        self.net = net or Network()
        n = self.net
        h = ctypes.c_int64(0)
        rc = _lib.lib().p3d_network_create(
            n.n_species, n.n_reactions, _lib.i32ptr(n.rtype), _lib.dptr(n.alpha),
            _lib.dptr(n.beta), _lib.dptr(n.gamma), _lib.i32ptr(n.reactants),
            _lib.i32ptr(n.products), int(n.e_index), _lib.i32ptr(n.roles),
            _lib.dptr(np.ascontiguousarray(n.post_w)), ctypes.byref(h))
        _lib.check(rc, "p3d_network_create")
        self.handle = h.value

    def close(self):
        if getattr(self, "handle", 0):
            _lib.lib().p3d_network_destroy(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def default_device_network():
    global _default
    if _default is None:
        _default = DeviceNetwork()
    return _default
