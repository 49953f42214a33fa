"""A larger synthetic network for the dense-LU path.

No network larger than the 31-species one ships with the reference, so the
path that needs at least 64 species (the FP64-MFMA trailing update of the
Newton matrix's LU) is exercised on a generated one: the bundled 31 species
and 75 reactions, plus 33 nitrogen- and hydrocarbon-bearing species with
reactions drawn from the usual gas-phase classes (dissociative
recombination, photodissociation and photoionisation, cosmic-ray
ionisation, ion-molecule, charge exchange, neutral-neutral), every one of
them balancing nuclei and charge.  Rate parameters are drawn from the
ranges those classes take in astrochemical databases, from a seeded
generator, so the network is reproducible; it is a workload, not a physical
model.
"""

import re

import numpy as np

from .network import REACTIONS, TRACKED

EXTRA = ['N', 'N+', 'NH', 'NH+', 'NH2', 'NH2+', 'NH3', 'NH3+', 'NH4+', 'N2', 'N2+', 'N2H+',
         'CN', 'CN+', 'HCN', 'HCN+', 'HCNH+', 'NO', 'NO+', 'HNO+',
         'CH2', 'CH3', 'CH3+', 'CH4', 'CH4+', 'CH5+', 'C2', 'C2+', 'C2H', 'C2H+', 'C2H2',
         'C2H2+', 'O2+']
# the further 39 species of the 103-species network: nitrogen oxides and
# isomers, sulphur, silicon and more carbon-oxygen species
EXTRA_WIDE = ['HNC', 'H2CN', 'HNO', 'NO2', 'N2O', 'OCN',
              'SH', 'SH+', 'H2S', 'H2S+', 'H3S+', 'CS', 'CS+', 'HCS+', 'SO', 'SO+', 'SO2', 'OCS',
              'OCS+', 'NS', 'NS+',
              'SiH', 'SiH+', 'SiO', 'SiO+', 'SiOH+', 'SiH2+',
              'HCO', 'HOC+', 'H2CO', 'H2CO+', 'H3CO+', 'CO2', 'HCO2+', 'O2H+', 'C3', 'C3H', 'C3H+',
              'C2H3+']

_TOKEN = re.compile(r'(He|Si|Fe|H|C|N|O|S)(\d*)')


def formula(species):
    """({element: count}, charge) of a species name such as 'C2H2+'."""
    if species == 'e-':
        return {}, -1
    body = species.rstrip('+-')
    charge = species.count('+') - species.count('-')
    atoms = {}
    pos = 0
    for m in _TOKEN.finditer(body):
        if m.start() != pos:
            raise ValueError(f"cannot parse {species!r}")
        atoms[m.group(1)] = atoms.get(m.group(1), 0) + int(m.group(2) or 1)
        pos = m.end()
    if pos != len(body):
        raise ValueError(f"cannot parse {species!r}")
    return atoms, charge


def _key(names):
    atoms, q = {}, 0
    for s in names:
        a, c = formula(s)
        for e, k in a.items():
            atoms[e] = atoms.get(e, 0) + k
        q += c
    return tuple(sorted(atoms.items())), q


def balanced(reactants, products):
    """Nuclei and charge balance of one reaction (the PAH- pseudo-species
    excluded, as the bundled network's PAH recombinations move charge out
    of the tracked species)."""
    r = [s for s in reactants if s != 'PAH-']
    return _key(r) == _key(products) if len(r) == len(reactants) else True


def synthetic_network(seed=2024, n_extra=None, max_channels=3, max_exchange=2, extra=None):
    """(species, reactions) of the 64-species synthetic network (31 bundled
    + 33 extra, or the first n_extra of the extra list, or the given extra
    species).  Reactions are the bundled 75 first, then the generated ones in
    a fixed order: up to max_channels product channels per recombination and
    max_exchange per two-body reactant pair."""
    if extra is None:
        extra = EXTRA if n_extra is None else EXTRA[:n_extra]
    species = list(TRACKED) + list(extra)
    rng = np.random.RandomState(seed)
    comp = {s: formula(s) for s in species}
    pairs = {}
    for i, a in enumerate(species):
        for b in species[i:]:
            pairs.setdefault(_key([a, b]), []).append((a, b))
    new = set(extra)
    out = []
    seen = set()

    def add(name, typ, reac, prod, a, b, g):
        sig = (tuple(sorted(reac)), tuple(sorted(prod)))
        if sig in seen or sorted(reac) == sorted(prod):
            return
        assert balanced(reac, prod), (reac, prod)
        seen.add(sig)
        out.append((name, typ, list(reac), list(prod), float(a), float(b), float(g)))

    for (_, _, reac, prod, *_r) in REACTIONS:
        seen.add((tuple(sorted(reac)), tuple(sorted(prod))))

    def neutral(x):
        return comp[x][1] == 0 and x != 'e-'

    def fragments(s, charge):
        """Two neutral fragments of species s (charge == its own) or of
        s + e- (a cation recombining), products in the species list."""
        key = _key([s]) if charge == comp[s][1] else _key([s, 'e-'])
        return [p for p in pairs.get(key, []) if s not in p and all(neutral(x) for x in p)]

    for s in species:
        if s not in new:
            continue
        atoms, q = comp[s]
        n_atoms = sum(atoms.values())
        if q == 1:
            # dissociative (or radiative) recombination
            ch = fragments(s, 0)
            rng.shuffle(ch)
            for p in ch[:max_channels]:
                add(f'{s}_DR_{"_".join(p)}', 'recomb', [s, 'e-'], list(p),
                    10 ** rng.uniform(-7.3, -6.5), -rng.uniform(0.3, 0.8), 0.0)
            if not ch and (s.rstrip('+') in comp):
                add(f'{s}_RR', 'recomb', [s, 'e-'], [s.rstrip('+')],
                    10 ** rng.uniform(-12.5, -11.5), -0.7, 0.0)
        else:
            if n_atoms > 1:
                ch = fragments(s, 0)
                rng.shuffle(ch)
                for p in ch[:2]:
                    add(f'{s}_photodiss_{"_".join(p)}', 'photodiss', [s], list(p),
                        10 ** rng.uniform(-10, -9), 0.0, rng.uniform(1.5, 3.0))
            ion = s + '+'
            if ion in comp:
                add(f'{s}_photoion', 'photoion', [s], [ion, 'e-'],
                    10 ** rng.uniform(-11, -9.5), 0.0, rng.uniform(2.0, 3.5))
                add(f'{s}_CR_ion', 'cosmic_ray', [s], [ion, 'e-'],
                    10 ** rng.uniform(-0.5, 0.7), 0.0, 0.0)
    # two-body exchanges with at least one new species on either side
    partners = ['H', 'H2', 'C', 'O', 'N', 'OH', 'H2O', 'CO', 'CH', 'NH', 'N2', 'H+', 'H3+',
                'C+', 'He+', 'O+', 'HCO+']
    partners = [p for p in partners if p in comp]
    for a in species:
        for b in partners:
            key = _key([a, b])
            qa, qb = comp[a][1], comp[b][1]
            if qa < 0 or qa + qb not in (0, 1):
                continue
            # products: neutrals, plus exactly one cation when a cation reacts
            cands = [p for p in pairs.get(key, [])
                     if sorted(p) != sorted((a, b)) and (new & ({a, b} | set(p)))
                     and all(comp[x][1] in (0, 1) and x != 'e-' for x in p)
                     and sum(comp[x][1] for x in p) == qa + qb]
            if not cands:
                continue
            rng.shuffle(cands)
            for p in cands[:max_exchange]:
                if qa + qb == 0:
                    typ, al, be, ga = 'two_body', 10 ** rng.uniform(-11.5, -10), \
                        rng.uniform(-0.3, 0.5), rng.choice([0.0, rng.uniform(100, 3000)])
                else:
                    ions_in = [x for x in (a, b) if comp[x][1] == 1]
                    ions_out = [x for x in p if comp[x][1] == 1]
                    ce = len(ions_in) == 1 and len(ions_out) == 1 and \
                        sorted([ions_in[0].rstrip('+'), ions_out[0]]) == \
                        sorted([ions_out[0].rstrip('+'), ions_in[0]])
                    typ = 'charge_exchange' if ce else 'two_body'
                    al, be, ga = 10 ** rng.uniform(-9.7, -8.8), 0.0, 0.0
                add(f'{a}_{b}_to_{"_".join(p)}', typ, [a, b], list(p), al, be, ga)
    reactions = list(REACTIONS) + out
    return species, reactions


def synthetic_network_wide(seed=2024):
    """(species, reactions) of the 103-species synthetic network (the 64 of
    synthetic_network plus EXTRA_WIDE), one product channel per two-body
    reactant pair: 1349 reactions, 3830 structural Jacobian nonzeros."""
    return synthetic_network(seed=seed, max_exchange=1, extra=EXTRA + EXTRA_WIDE)


def element_totals_check(species, reactions):
    """Indices of reactions that do not balance nuclei and charge."""
    return [r for r, (_, _, reac, prod, *_x) in enumerate(reactions) if not balanced(reac, prod)]
