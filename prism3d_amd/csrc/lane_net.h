// Generated by prism3d_amd/gen_lane.py from the network structure of
// prism3d_amd/network.py.  Do not edit; rerun the generator.
//
// 31 species, 75 reactions, 232 structural Jacobian entries,
// 260 entries in the filled no-pivot LU factors.
#pragma once

namespace p3d {
namespace lane {

constexpr int kNS = 31, kNR = 75, kNJ = 232, kNLU = 260, kE = 0;
constexpr int kRoles[7] = {1, 2, 3, 7, 8, 9, 13};
constexpr int kRtype[kNR] = {5, 2, 4, 4, 0, 0, 0, 1, 0, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 1, 0, 1, 0, 0, 0, 4, 0, 0, 0, 0, 0, 0, 0, 4, 4, 4, 4, 0, 0, 1, 4, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
constexpr int kReact[2 * kNR] = {1, 1, 2, -1, 2, -1, 1, -1, 3, 0, 1, 0, 4, 1, 8, -1, 7, 0, 9, -1, 7, 15, 7, 17, 8, 15, 10, 13, 7, 2, 11, 2, 11, 0, 13, 3, 14, 1, 13, 6, 16, 2, 18, 2, 19, 0, 19, 0, 15, -1, 17, -1, 13, 2, 15, 1, 15, 2, 9, 6, 22, 0, 22, -1, 5, 2, 6, 0, 6, 0, 6, 8, 23, -1, 24, 0, 25, -1, 26, 0, 27, -1, 28, 0, 21, 1, 21, 2, 29, -1, 30, 9, 30, 0, 7, -2, 24, -2, 3, -2, 26, -2, 28, -2, 9, -1, 17, -1, 15, -1, 8, -1, 13, 15, 8, 20, 20, -1, 20, -1, 20, 7, 10, 1, 10, -1, 12, 0, 12, 0, 10, 13, 21, 0, 8, 17, 2, 1, 2, 2, 28, 0, 26, 0, 16, 0, 18, 0, 18, 0};
constexpr int kProd[3 * kNR] = {2, -1, -1, 1, 1, -1, 5, 0, -1, 3, 0, -1, 1, -1, -1, 4, -1, -1, 2, 0, -1, 7, 0, -1, 8, -1, -1, 8, 13, -1, 9, 3, -1, 22, 1, -1, 9, 1, -1, 9, 1, -1, 11, 1, -1, 12, 1, -1, 8, 1, -1, 14, 1, -1, 13, 3, -1, 16, 2, -1, 18, 1, -1, 19, 1, -1, 15, 1, 1, 17, 1, -1, 13, 1, -1, 15, 1, -1, 15, 1, -1, 13, 2, -1, 17, 1, -1, 22, 2, -1, 9, 1, -1, 21, 1, -1, 6, 1, -1, 2, 1, -1, 1, 1, 1, 11, 2, -1, 24, 0, -1, 23, -1, -1, 26, 0, -1, 25, -1, -1, 28, 0, -1, 27, -1, -1, 9, 3, -1, 22, 1, -1, 30, 0, -1, 7, 13, 29, 29, -1, -1, 8, -1, -1, 23, -1, -1, 1, -1, -1, 25, -1, -1, 27, -1, -1, 8, 13, -1, 15, 1, -1, 13, 1, -1, 7, 0, -1, 20, 1, -1, 9, 13, -1, 13, 13, -1, 13, 13, -1, 21, 13, -1, 8, 2, -1, 8, 1, -1, 10, 1, -1, 8, 1, 1, 22, 0, -1, 8, 13, -1, 9, 1, 1, 1, 1, 1, 2, 1, 1, 27, -1, -1, 25, -1, -1, 13, 1, -1, 13, 1, 1, 15, 1, -1};
// elimination order: position -> species
constexpr int kPerm[kNS] = {5, 23, 24, 25, 26, 27, 28, 4, 29, 30, 14, 12, 19, 3, 18, 20, 16, 11, 22, 10, 6, 17, 7, 15, 0, 1, 2, 8, 9, 13, 21};

// dy/dt (chemistry/solver.py:69-143) at the abundances held in slot
// entries yoff..yoff+kNS, floored at 1e-30 as the rates see them;
// accumulation per species follows the reaction order.  The abundances
// are read once into registers; rate coefficients are read per reaction.
template <class Slot>
P3D_HD P3D_INLINE void rhs(const Cell &p, const Slot &m, int yoff, double (&f)[kNS]) {
#pragma unroll
  for (int i = 0; i < kNS; ++i) f[i] = 0.0;
  double yv[kNS];
#pragma unroll
  for (int i = 0; i < kNS; ++i) yv[i] = fmax((double)m.at(yoff + i), 1e-30);
  double pm, dpm;
  pah_minus(yv[kE], p, pm, dpm);
  (void)dpm;
  { const double ya = yv[1]; const double q = m.k(0) * ya * p.nH * 0.5; f[1] -= q; f[1] -= q; f[2] += q; }  // H2_formation_grain
  { const double ya = yv[2]; const double q = m.k(1) * ya; f[2] -= q; f[1] += q; f[1] += q; }  // H2_photodiss
  { const double ya = yv[2]; const double q = m.k(2) * ya; f[2] -= q; f[5] += q; f[0] += q; }  // H2_CR_ionization
  { const double ya = yv[1]; const double q = m.k(3) * ya; f[1] -= q; f[3] += q; f[0] += q; }  // H_CR_ionization
  { const double ya = yv[3]; const double yb = yv[0]; const double q = m.k(4) * ya * yb * p.nH; f[3] -= q; f[0] -= q; f[1] += q; }  // H_recombination
  { const double ya = yv[1]; const double yb = yv[0]; const double q = m.k(5) * ya * yb * p.nH; f[1] -= q; f[0] -= q; f[4] += q; }  // H_minus_formation
  { const double ya = yv[4]; const double yb = yv[1]; const double q = m.k(6) * ya * yb * p.nH; f[4] -= q; f[1] -= q; f[2] += q; f[0] += q; }  // H2_gasphase_Hminus
  { const double ya = yv[8]; const double q = m.k(7) * ya; f[8] -= q; f[7] += q; f[0] += q; }  // C_photoionization
  { const double ya = yv[7]; const double yb = yv[0]; const double q = m.k(8) * ya * yb * p.nH; f[7] -= q; f[0] -= q; f[8] += q; }  // Cp_recombination
  { const double ya = yv[9]; const double q = m.k(9) * ya; f[9] -= q; f[8] += q; f[13] += q; }  // CO_photodiss
  { const double ya = yv[7]; const double yb = yv[15]; const double q = m.k(10) * ya * yb * p.nH; f[7] -= q; f[15] -= q; f[9] += q; f[3] += q; }  // CO_form_CpOH
  { const double ya = yv[7]; const double yb = yv[17]; const double q = m.k(11) * ya * yb * p.nH; f[7] -= q; f[17] -= q; f[22] += q; f[1] += q; }  // HCOp_form_CpH2O
  { const double ya = yv[8]; const double yb = yv[15]; const double q = m.k(12) * ya * yb * p.nH; f[8] -= q; f[15] -= q; f[9] += q; f[1] += q; }  // CO_form_COH
  { const double ya = yv[10]; const double yb = yv[13]; const double q = m.k(13) * ya * yb * p.nH; f[10] -= q; f[13] -= q; f[9] += q; f[1] += q; }  // CO_form_CHO
  { const double ya = yv[7]; const double yb = yv[2]; const double q = m.k(14) * ya * yb * p.nH; f[7] -= q; f[2] -= q; f[11] += q; f[1] += q; }  // CHp_form
  { const double ya = yv[11]; const double yb = yv[2]; const double q = m.k(15) * ya * yb * p.nH; f[11] -= q; f[2] -= q; f[12] += q; f[1] += q; }  // CH2p_form
  { const double ya = yv[11]; const double yb = yv[0]; const double q = m.k(16) * ya * yb * p.nH; f[11] -= q; f[0] -= q; f[8] += q; f[1] += q; }  // CHp_recomb
  { const double ya = yv[13]; const double yb = yv[3]; const double q = m.k(17) * ya * yb * p.nH; f[13] -= q; f[3] -= q; f[14] += q; f[1] += q; }  // O_Hp_CE
  { const double ya = yv[14]; const double yb = yv[1]; const double q = m.k(18) * ya * yb * p.nH; f[14] -= q; f[1] -= q; f[13] += q; f[3] += q; }  // Op_H_CE
  { const double ya = yv[13]; const double yb = yv[6]; const double q = m.k(19) * ya * yb * p.nH; f[13] -= q; f[6] -= q; f[16] += q; f[2] += q; }  // OHp_form_OH3p
  { const double ya = yv[16]; const double yb = yv[2]; const double q = m.k(20) * ya * yb * p.nH; f[16] -= q; f[2] -= q; f[18] += q; f[1] += q; }  // H2Op_form
  { const double ya = yv[18]; const double yb = yv[2]; const double q = m.k(21) * ya * yb * p.nH; f[18] -= q; f[2] -= q; f[19] += q; f[1] += q; }  // H3Op_form
  { const double ya = yv[19]; const double yb = yv[0]; const double q = m.k(22) * ya * yb * p.nH; f[19] -= q; f[0] -= q; f[15] += q; f[1] += q; f[1] += q; }  // H3Op_recomb_OH
  { const double ya = yv[19]; const double yb = yv[0]; const double q = m.k(23) * ya * yb * p.nH; f[19] -= q; f[0] -= q; f[17] += q; f[1] += q; }  // H3Op_recomb_H2O
  { const double ya = yv[15]; const double q = m.k(24) * ya; f[15] -= q; f[13] += q; f[1] += q; }  // OH_photodiss
  { const double ya = yv[17]; const double q = m.k(25) * ya; f[17] -= q; f[15] += q; f[1] += q; }  // H2O_photodiss
  { const double ya = yv[13]; const double yb = yv[2]; const double q = m.k(26) * ya * yb * p.nH; f[13] -= q; f[2] -= q; f[15] += q; f[1] += q; }  // OH_form_OH2
  { const double ya = yv[15]; const double yb = yv[1]; const double q = m.k(27) * ya * yb * p.nH; f[15] -= q; f[1] -= q; f[13] += q; f[2] += q; }  // OH_destr_OHH
  { const double ya = yv[15]; const double yb = yv[2]; const double q = m.k(28) * ya * yb * p.nH; f[15] -= q; f[2] -= q; f[17] += q; f[1] += q; }  // H2O_form_OHH2
  { const double ya = yv[9]; const double yb = yv[6]; const double q = m.k(29) * ya * yb * p.nH; f[9] -= q; f[6] -= q; f[22] += q; f[2] += q; }  // HCOp_form_COH3p
  { const double ya = yv[22]; const double yb = yv[0]; const double q = m.k(30) * ya * yb * p.nH; f[22] -= q; f[0] -= q; f[9] += q; f[1] += q; }  // HCOp_recomb
  { const double ya = yv[22]; const double q = m.k(31) * ya; f[22] -= q; f[21] += q; f[1] += q; }  // HCOp_photodiss
  { const double ya = yv[5]; const double yb = yv[2]; const double q = m.k(32) * ya * yb * p.nH; f[5] -= q; f[2] -= q; f[6] += q; f[1] += q; }  // H3p_form
  { const double ya = yv[6]; const double yb = yv[0]; const double q = m.k(33) * ya * yb * p.nH; f[6] -= q; f[0] -= q; f[2] += q; f[1] += q; }  // H3p_recomb_H2
  { const double ya = yv[6]; const double yb = yv[0]; const double q = m.k(34) * ya * yb * p.nH; f[6] -= q; f[0] -= q; f[1] += q; f[1] += q; f[1] += q; }  // H3p_recomb_3H
  { const double ya = yv[6]; const double yb = yv[8]; const double q = m.k(35) * ya * yb * p.nH; f[6] -= q; f[8] -= q; f[11] += q; f[2] += q; }  // CHp_form_H3pC
  { const double ya = yv[23]; const double q = m.k(36) * ya; f[23] -= q; f[24] += q; f[0] += q; }  // S_photoion
  { const double ya = yv[24]; const double yb = yv[0]; const double q = m.k(37) * ya * yb * p.nH; f[24] -= q; f[0] -= q; f[23] += q; }  // Sp_recomb
  { const double ya = yv[25]; const double q = m.k(38) * ya; f[25] -= q; f[26] += q; f[0] += q; }  // Si_photoion
  { const double ya = yv[26]; const double yb = yv[0]; const double q = m.k(39) * ya * yb * p.nH; f[26] -= q; f[0] -= q; f[25] += q; }  // Sip_recomb
  { const double ya = yv[27]; const double q = m.k(40) * ya; f[27] -= q; f[28] += q; f[0] += q; }  // Fe_photoion
  { const double ya = yv[28]; const double yb = yv[0]; const double q = m.k(41) * ya * yb * p.nH; f[28] -= q; f[0] -= q; f[27] += q; }  // Fep_recomb
  { const double ya = yv[21]; const double yb = yv[1]; const double q = m.k(42) * ya * yb * p.nH; f[21] -= q; f[1] -= q; f[9] += q; f[3] += q; }  // COp_H
  { const double ya = yv[21]; const double yb = yv[2]; const double q = m.k(43) * ya * yb * p.nH; f[21] -= q; f[2] -= q; f[22] += q; f[1] += q; }  // COp_H2_HCOp
  { const double ya = yv[29]; const double q = m.k(44) * ya; f[29] -= q; f[30] += q; f[0] += q; }  // He_CR_ion
  { const double ya = yv[30]; const double yb = yv[9]; const double q = m.k(45) * ya * yb * p.nH; f[30] -= q; f[9] -= q; f[7] += q; f[13] += q; f[29] += q; }  // Hep_CO_destruct
  { const double ya = yv[30]; const double yb = yv[0]; const double q = m.k(46) * ya * yb * p.nH; f[30] -= q; f[0] -= q; f[29] += q; }  // Hep_recomb
  { const double ya = yv[7]; const double q = m.k(47) * ya * pm * p.nH; f[7] -= q; f[8] += q; }  // Cp_PAH_recomb
  { const double ya = yv[24]; const double q = m.k(48) * ya * pm * p.nH; f[24] -= q; f[23] += q; }  // Sp_PAH_recomb
  { const double ya = yv[3]; const double q = m.k(49) * ya * pm * p.nH; f[3] -= q; f[1] += q; }  // Hp_PAH_recomb
  { const double ya = yv[26]; const double q = m.k(50) * ya * pm * p.nH; f[26] -= q; f[25] += q; }  // Sip_PAH_recomb
  { const double ya = yv[28]; const double q = m.k(51) * ya * pm * p.nH; f[28] -= q; f[27] += q; }  // Fep_PAH_recomb
  { const double ya = yv[9]; const double q = m.k(52) * ya; f[9] -= q; f[8] += q; f[13] += q; }  // CO_CR_photodiss
  { const double ya = yv[17]; const double q = m.k(53) * ya; f[17] -= q; f[15] += q; f[1] += q; }  // H2O_CR_photodiss
  { const double ya = yv[15]; const double q = m.k(54) * ya; f[15] -= q; f[13] += q; f[1] += q; }  // OH_CR_photodiss
  { const double ya = yv[8]; const double q = m.k(55) * ya; f[8] -= q; f[7] += q; f[0] += q; }  // C_CR_photoion
  { const double ya = yv[13]; const double yb = yv[15]; const double q = m.k(56) * ya * yb * p.nH; f[13] -= q; f[15] -= q; f[20] += q; f[1] += q; }  // O2_form_OOH
  { const double ya = yv[8]; const double yb = yv[20]; const double q = m.k(57) * ya * yb * p.nH; f[8] -= q; f[20] -= q; f[9] += q; f[13] += q; }  // CO_form_CO2
  { const double ya = yv[20]; const double q = m.k(58) * ya; f[20] -= q; f[13] += q; f[13] += q; }  // O2_photodiss
  { const double ya = yv[20]; const double q = m.k(59) * ya; f[20] -= q; f[13] += q; f[13] += q; }  // O2_CR_photodiss
  { const double ya = yv[20]; const double yb = yv[7]; const double q = m.k(60) * ya * yb * p.nH; f[20] -= q; f[7] -= q; f[21] += q; f[13] += q; }  // COp_form_O2Cp
  { const double ya = yv[10]; const double yb = yv[1]; const double q = m.k(61) * ya * yb * p.nH; f[10] -= q; f[1] -= q; f[8] += q; f[2] += q; }  // CH_destr_H
  { const double ya = yv[10]; const double q = m.k(62) * ya; f[10] -= q; f[8] += q; f[1] += q; }  // CH_photodiss
  { const double ya = yv[12]; const double yb = yv[0]; const double q = m.k(63) * ya * yb * p.nH; f[12] -= q; f[0] -= q; f[10] += q; f[1] += q; }  // CH2p_recomb_CH
  { const double ya = yv[12]; const double yb = yv[0]; const double q = m.k(64) * ya * yb * p.nH; f[12] -= q; f[0] -= q; f[8] += q; f[1] += q; f[1] += q; }  // CH2p_recomb_C
  { const double ya = yv[10]; const double yb = yv[13]; const double q = m.k(65) * ya * yb * p.nH; f[10] -= q; f[13] -= q; f[22] += q; f[0] += q; }  // HCOp_form_CHO
  { const double ya = yv[21]; const double yb = yv[0]; const double q = m.k(66) * ya * yb * p.nH; f[21] -= q; f[0] -= q; f[8] += q; f[13] += q; }  // COp_recomb
  { const double ya = yv[8]; const double yb = yv[17]; const double q = m.k(67) * ya * yb * p.nH; f[8] -= q; f[17] -= q; f[9] += q; f[1] += q; f[1] += q; }  // CO_form_CH2O
  { const double ya = yv[2]; const double yb = yv[1]; const double q = m.k(68) * ya * yb * p.nH; f[2] -= q; f[1] -= q; f[1] += q; f[1] += q; f[1] += q; }  // H2_coll_dissoc_H
  { const double ya = yv[2]; const double q = m.k(69) * ya * ya * p.nH; f[2] -= q; f[2] -= q; f[2] += q; f[1] += q; f[1] += q; }  // H2_coll_dissoc_H2
  { const double ya = yv[28]; const double yb = yv[0]; const double q = m.k(70) * ya * yb * p.nH; f[28] -= q; f[0] -= q; f[27] += q; }  // Fep_grain_recomb
  { const double ya = yv[26]; const double yb = yv[0]; const double q = m.k(71) * ya * yb * p.nH; f[26] -= q; f[0] -= q; f[25] += q; }  // Sip_grain_recomb
  { const double ya = yv[16]; const double yb = yv[0]; const double q = m.k(72) * ya * yb * p.nH; f[16] -= q; f[0] -= q; f[13] += q; f[1] += q; }  // OHp_recomb
  { const double ya = yv[18]; const double yb = yv[0]; const double q = m.k(73) * ya * yb * p.nH; f[18] -= q; f[0] -= q; f[13] += q; f[1] += q; f[1] += q; }  // H2Op_recomb_O
  { const double ya = yv[18]; const double yb = yv[0]; const double q = m.k(74) * ya * yb * p.nH; f[18] -= q; f[0] -= q; f[15] += q; f[1] += q; }  // H2Op_recomb_OH
}

// Analytic Jacobian at the abundances held in slot entries yoff..yoff+kNS
// (floored at 1e-30 as the rates see them) into the slot's J entries, one
// row at a time; abundances are read per term so that only the current
// row's entries stay in registers.
template <class Slot>
P3D_HD P3D_INLINE void jacobian(const Cell &p, const Slot &m, int yoff) {
  const double ye = fmax((double)m.at(yoff + kE), 1e-30);
  double pm, dpm;
  pah_minus(ye, p, pm, dpm);
  {  // rows H2+, S
    double j0_0 = 0.0;
    double j0_26 = 0.0;
    double j1_1 = 0.0;
    double j1_2 = 0.0;
    double j1_24 = 0.0;
    { const double k = m.k(2); const double ya = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j0_26 += 1.0 * d; }
    { const double k = m.k(32); const double ya = fmax((double)m.at(yoff + 5), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j0_0 += -1.0 * d; }
    { const double k = m.k(32); const double ya = fmax((double)m.at(yoff + 5), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j0_26 += -1.0 * d; }
    { const double k = m.k(36); const double ya = fmax((double)m.at(yoff + 23), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j1_1 += -1.0 * d; }
    { const double k = m.k(37); const double ya = fmax((double)m.at(yoff + 24), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j1_2 += 1.0 * d; }
    { const double k = m.k(37); const double ya = fmax((double)m.at(yoff + 24), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j1_24 += 1.0 * d; }
    { const double k = m.k(48); const double ya = fmax((double)m.at(yoff + 24), 1e-30); const double d = k * pm * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j1_2 += 1.0 * d; }
    { const double k = m.k(48); const double ya = fmax((double)m.at(yoff + 24), 1e-30); const double d = k * ya * p.nH * dpm * (ye > 1e-30 ? 1.0 : 0.0); j1_24 += 1.0 * d; }
    P3D_FENCE();
    m.J(0) = j0_0;
    m.J(1) = j0_26;
    m.J(2) = j1_1;
    m.J(3) = j1_2;
    m.J(4) = j1_24;
  }
  P3D_FENCE();
  {  // rows S+, Si
    double j2_1 = 0.0;
    double j2_2 = 0.0;
    double j2_24 = 0.0;
    double j3_3 = 0.0;
    double j3_4 = 0.0;
    double j3_24 = 0.0;
    { const double k = m.k(36); const double ya = fmax((double)m.at(yoff + 23), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j2_1 += 1.0 * d; }
    { const double k = m.k(37); const double ya = fmax((double)m.at(yoff + 24), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j2_2 += -1.0 * d; }
    { const double k = m.k(37); const double ya = fmax((double)m.at(yoff + 24), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j2_24 += -1.0 * d; }
    { const double k = m.k(48); const double ya = fmax((double)m.at(yoff + 24), 1e-30); const double d = k * pm * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j2_2 += -1.0 * d; }
    { const double k = m.k(48); const double ya = fmax((double)m.at(yoff + 24), 1e-30); const double d = k * ya * p.nH * dpm * (ye > 1e-30 ? 1.0 : 0.0); j2_24 += -1.0 * d; }
    { const double k = m.k(38); const double ya = fmax((double)m.at(yoff + 25), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j3_3 += -1.0 * d; }
    { const double k = m.k(39); const double ya = fmax((double)m.at(yoff + 26), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j3_4 += 1.0 * d; }
    { const double k = m.k(39); const double ya = fmax((double)m.at(yoff + 26), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j3_24 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(50); const double ya = fmax((double)m.at(yoff + 26), 1e-30); const double d = k * pm * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j3_4 += 1.0 * d; }
    { const double k = m.k(50); const double ya = fmax((double)m.at(yoff + 26), 1e-30); const double d = k * ya * p.nH * dpm * (ye > 1e-30 ? 1.0 : 0.0); j3_24 += 1.0 * d; }
    { const double k = m.k(71); const double ya = fmax((double)m.at(yoff + 26), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j3_4 += 1.0 * d; }
    { const double k = m.k(71); const double ya = fmax((double)m.at(yoff + 26), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j3_24 += 1.0 * d; }
    m.J(5) = j2_1;
    m.J(6) = j2_2;
    m.J(7) = j2_24;
    m.J(8) = j3_3;
    m.J(9) = j3_4;
    m.J(10) = j3_24;
  }
  P3D_FENCE();
  {  // rows Si+, Fe
    double j4_3 = 0.0;
    double j4_4 = 0.0;
    double j4_24 = 0.0;
    double j5_5 = 0.0;
    double j5_6 = 0.0;
    double j5_24 = 0.0;
    { const double k = m.k(38); const double ya = fmax((double)m.at(yoff + 25), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j4_3 += 1.0 * d; }
    { const double k = m.k(39); const double ya = fmax((double)m.at(yoff + 26), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j4_4 += -1.0 * d; }
    { const double k = m.k(39); const double ya = fmax((double)m.at(yoff + 26), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j4_24 += -1.0 * d; }
    { const double k = m.k(50); const double ya = fmax((double)m.at(yoff + 26), 1e-30); const double d = k * pm * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j4_4 += -1.0 * d; }
    { const double k = m.k(50); const double ya = fmax((double)m.at(yoff + 26), 1e-30); const double d = k * ya * p.nH * dpm * (ye > 1e-30 ? 1.0 : 0.0); j4_24 += -1.0 * d; }
    { const double k = m.k(71); const double ya = fmax((double)m.at(yoff + 26), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j4_4 += -1.0 * d; }
    { const double k = m.k(71); const double ya = fmax((double)m.at(yoff + 26), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j4_24 += -1.0 * d; }
    { const double k = m.k(40); const double ya = fmax((double)m.at(yoff + 27), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j5_5 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(41); const double ya = fmax((double)m.at(yoff + 28), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j5_6 += 1.0 * d; }
    { const double k = m.k(41); const double ya = fmax((double)m.at(yoff + 28), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j5_24 += 1.0 * d; }
    { const double k = m.k(51); const double ya = fmax((double)m.at(yoff + 28), 1e-30); const double d = k * pm * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j5_6 += 1.0 * d; }
    { const double k = m.k(51); const double ya = fmax((double)m.at(yoff + 28), 1e-30); const double d = k * ya * p.nH * dpm * (ye > 1e-30 ? 1.0 : 0.0); j5_24 += 1.0 * d; }
    { const double k = m.k(70); const double ya = fmax((double)m.at(yoff + 28), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j5_6 += 1.0 * d; }
    { const double k = m.k(70); const double ya = fmax((double)m.at(yoff + 28), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j5_24 += 1.0 * d; }
    m.J(11) = j4_3;
    m.J(12) = j4_4;
    m.J(13) = j4_24;
    m.J(14) = j5_5;
    m.J(15) = j5_6;
    m.J(16) = j5_24;
  }
  P3D_FENCE();
  {  // rows Fe+, H-
    double j6_5 = 0.0;
    double j6_6 = 0.0;
    double j6_24 = 0.0;
    double j7_7 = 0.0;
    double j7_24 = 0.0;
    double j7_25 = 0.0;
    { const double k = m.k(40); const double ya = fmax((double)m.at(yoff + 27), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j6_5 += 1.0 * d; }
    { const double k = m.k(41); const double ya = fmax((double)m.at(yoff + 28), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j6_6 += -1.0 * d; }
    { const double k = m.k(41); const double ya = fmax((double)m.at(yoff + 28), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j6_24 += -1.0 * d; }
    { const double k = m.k(51); const double ya = fmax((double)m.at(yoff + 28), 1e-30); const double d = k * pm * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j6_6 += -1.0 * d; }
    { const double k = m.k(51); const double ya = fmax((double)m.at(yoff + 28), 1e-30); const double d = k * ya * p.nH * dpm * (ye > 1e-30 ? 1.0 : 0.0); j6_24 += -1.0 * d; }
    { const double k = m.k(70); const double ya = fmax((double)m.at(yoff + 28), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j6_6 += -1.0 * d; }
    { const double k = m.k(70); const double ya = fmax((double)m.at(yoff + 28), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j6_24 += -1.0 * d; }
    { const double k = m.k(5); const double ya = fmax((double)m.at(yoff + 1), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j7_25 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(5); const double ya = fmax((double)m.at(yoff + 1), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j7_24 += 1.0 * d; }
    { const double k = m.k(6); const double ya = fmax((double)m.at(yoff + 4), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j7_7 += -1.0 * d; }
    { const double k = m.k(6); const double ya = fmax((double)m.at(yoff + 4), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j7_25 += -1.0 * d; }
    m.J(17) = j6_5;
    m.J(18) = j6_6;
    m.J(19) = j6_24;
    m.J(20) = j7_7;
    m.J(21) = j7_24;
    m.J(22) = j7_25;
  }
  P3D_FENCE();
  {  // rows He, He+
    double j8_8 = 0.0;
    double j8_9 = 0.0;
    double j8_24 = 0.0;
    double j8_28 = 0.0;
    double j9_8 = 0.0;
    double j9_9 = 0.0;
    double j9_24 = 0.0;
    double j9_28 = 0.0;
    { const double k = m.k(44); const double ya = fmax((double)m.at(yoff + 29), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j8_8 += -1.0 * d; }
    { const double k = m.k(45); const double ya = fmax((double)m.at(yoff + 30), 1e-30); const double yb = fmax((double)m.at(yoff + 9), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j8_9 += 1.0 * d; }
    { const double k = m.k(45); const double ya = fmax((double)m.at(yoff + 30), 1e-30); const double yb = fmax((double)m.at(yoff + 9), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j8_28 += 1.0 * d; }
    { const double k = m.k(46); const double ya = fmax((double)m.at(yoff + 30), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j8_9 += 1.0 * d; }
    { const double k = m.k(46); const double ya = fmax((double)m.at(yoff + 30), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j8_24 += 1.0 * d; }
    { const double k = m.k(44); const double ya = fmax((double)m.at(yoff + 29), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j9_8 += 1.0 * d; }
    { const double k = m.k(45); const double ya = fmax((double)m.at(yoff + 30), 1e-30); const double yb = fmax((double)m.at(yoff + 9), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j9_9 += -1.0 * d; }
    { const double k = m.k(45); const double ya = fmax((double)m.at(yoff + 30), 1e-30); const double yb = fmax((double)m.at(yoff + 9), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j9_28 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(46); const double ya = fmax((double)m.at(yoff + 30), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j9_9 += -1.0 * d; }
    { const double k = m.k(46); const double ya = fmax((double)m.at(yoff + 30), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j9_24 += -1.0 * d; }
    m.J(23) = j8_8;
    m.J(24) = j8_9;
    m.J(25) = j8_24;
    m.J(26) = j8_28;
    m.J(27) = j9_8;
    m.J(28) = j9_9;
    m.J(29) = j9_24;
    m.J(30) = j9_28;
  }
  P3D_FENCE();
  {  // rows O+, CH2+
    double j10_10 = 0.0;
    double j10_13 = 0.0;
    double j10_25 = 0.0;
    double j10_29 = 0.0;
    double j11_11 = 0.0;
    double j11_17 = 0.0;
    double j11_24 = 0.0;
    double j11_26 = 0.0;
    { const double k = m.k(17); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 3), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j10_29 += 1.0 * d; }
    { const double k = m.k(17); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 3), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j10_13 += 1.0 * d; }
    { const double k = m.k(18); const double ya = fmax((double)m.at(yoff + 14), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j10_10 += -1.0 * d; }
    { const double k = m.k(18); const double ya = fmax((double)m.at(yoff + 14), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j10_25 += -1.0 * d; }
    { const double k = m.k(15); const double ya = fmax((double)m.at(yoff + 11), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j11_17 += 1.0 * d; }
    { const double k = m.k(15); const double ya = fmax((double)m.at(yoff + 11), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j11_26 += 1.0 * d; }
    { const double k = m.k(63); const double ya = fmax((double)m.at(yoff + 12), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j11_11 += -1.0 * d; }
    { const double k = m.k(63); const double ya = fmax((double)m.at(yoff + 12), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j11_24 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(64); const double ya = fmax((double)m.at(yoff + 12), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j11_11 += -1.0 * d; }
    { const double k = m.k(64); const double ya = fmax((double)m.at(yoff + 12), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j11_24 += -1.0 * d; }
    m.J(31) = j10_10;
    m.J(32) = j10_13;
    m.J(33) = j10_25;
    m.J(34) = j10_29;
    m.J(35) = j11_11;
    m.J(36) = j11_17;
    m.J(37) = j11_24;
    m.J(38) = j11_26;
  }
  P3D_FENCE();
  {  // rows H3O+, H+
    double j12_12 = 0.0;
    double j12_14 = 0.0;
    double j12_24 = 0.0;
    double j12_26 = 0.0;
    double j13_10 = 0.0;
    double j13_13 = 0.0;
    double j13_22 = 0.0;
    double j13_23 = 0.0;
    double j13_24 = 0.0;
    double j13_25 = 0.0;
    double j13_29 = 0.0;
    double j13_30 = 0.0;
    { const double k = m.k(21); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j12_14 += 1.0 * d; }
    { const double k = m.k(21); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j12_26 += 1.0 * d; }
    { const double k = m.k(22); const double ya = fmax((double)m.at(yoff + 19), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j12_12 += -1.0 * d; }
    { const double k = m.k(22); const double ya = fmax((double)m.at(yoff + 19), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j12_24 += -1.0 * d; }
    { const double k = m.k(23); const double ya = fmax((double)m.at(yoff + 19), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j12_12 += -1.0 * d; }
    { const double k = m.k(23); const double ya = fmax((double)m.at(yoff + 19), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j12_24 += -1.0 * d; }
    { const double k = m.k(3); const double ya = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j13_25 += 1.0 * d; }
    { const double k = m.k(4); const double ya = fmax((double)m.at(yoff + 3), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j13_13 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(4); const double ya = fmax((double)m.at(yoff + 3), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j13_24 += -1.0 * d; }
    { const double k = m.k(10); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j13_22 += 1.0 * d; }
    { const double k = m.k(10); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j13_23 += 1.0 * d; }
    { const double k = m.k(17); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 3), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j13_29 += -1.0 * d; }
    { const double k = m.k(17); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 3), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j13_13 += -1.0 * d; }
    { const double k = m.k(18); const double ya = fmax((double)m.at(yoff + 14), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j13_10 += 1.0 * d; }
    { const double k = m.k(18); const double ya = fmax((double)m.at(yoff + 14), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j13_25 += 1.0 * d; }
    { const double k = m.k(42); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j13_30 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(42); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j13_25 += 1.0 * d; }
    { const double k = m.k(49); const double ya = fmax((double)m.at(yoff + 3), 1e-30); const double d = k * pm * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j13_13 += -1.0 * d; }
    { const double k = m.k(49); const double ya = fmax((double)m.at(yoff + 3), 1e-30); const double d = k * ya * p.nH * dpm * (ye > 1e-30 ? 1.0 : 0.0); j13_24 += -1.0 * d; }
    m.J(39) = j12_12;
    m.J(40) = j12_14;
    m.J(41) = j12_24;
    m.J(42) = j12_26;
    m.J(43) = j13_10;
    m.J(44) = j13_13;
    m.J(45) = j13_22;
    m.J(46) = j13_23;
    m.J(47) = j13_24;
    m.J(48) = j13_25;
    m.J(49) = j13_29;
    m.J(50) = j13_30;
  }
  P3D_FENCE();
  {  // rows H2O+, O2
    double j14_14 = 0.0;
    double j14_16 = 0.0;
    double j14_24 = 0.0;
    double j14_26 = 0.0;
    double j15_15 = 0.0;
    double j15_22 = 0.0;
    double j15_23 = 0.0;
    double j15_27 = 0.0;
    double j15_29 = 0.0;
    { const double k = m.k(20); const double ya = fmax((double)m.at(yoff + 16), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j14_16 += 1.0 * d; }
    { const double k = m.k(20); const double ya = fmax((double)m.at(yoff + 16), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j14_26 += 1.0 * d; }
    { const double k = m.k(21); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j14_14 += -1.0 * d; }
    { const double k = m.k(21); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j14_26 += -1.0 * d; }
    { const double k = m.k(73); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j14_14 += -1.0 * d; }
    { const double k = m.k(73); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j14_24 += -1.0 * d; }
    { const double k = m.k(74); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j14_14 += -1.0 * d; }
    { const double k = m.k(74); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j14_24 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(56); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j15_29 += 1.0 * d; }
    { const double k = m.k(56); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j15_23 += 1.0 * d; }
    { const double k = m.k(57); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 20), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j15_27 += -1.0 * d; }
    { const double k = m.k(57); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 20), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j15_15 += -1.0 * d; }
    { const double k = m.k(58); const double ya = fmax((double)m.at(yoff + 20), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j15_15 += -1.0 * d; }
    { const double k = m.k(59); const double ya = fmax((double)m.at(yoff + 20), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j15_15 += -1.0 * d; }
    { const double k = m.k(60); const double ya = fmax((double)m.at(yoff + 20), 1e-30); const double yb = fmax((double)m.at(yoff + 7), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j15_15 += -1.0 * d; }
    { const double k = m.k(60); const double ya = fmax((double)m.at(yoff + 20), 1e-30); const double yb = fmax((double)m.at(yoff + 7), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j15_22 += -1.0 * d; }
    P3D_FENCE();
    m.J(51) = j14_14;
    m.J(52) = j14_16;
    m.J(53) = j14_24;
    m.J(54) = j14_26;
    m.J(55) = j15_15;
    m.J(56) = j15_22;
    m.J(57) = j15_23;
    m.J(58) = j15_27;
    m.J(59) = j15_29;
  }
  P3D_FENCE();
  {  // rows OH+, CH+
    double j16_16 = 0.0;
    double j16_20 = 0.0;
    double j16_24 = 0.0;
    double j16_26 = 0.0;
    double j16_29 = 0.0;
    double j17_17 = 0.0;
    double j17_20 = 0.0;
    double j17_22 = 0.0;
    double j17_24 = 0.0;
    double j17_26 = 0.0;
    double j17_27 = 0.0;
    { const double k = m.k(19); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 6), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j16_29 += 1.0 * d; }
    { const double k = m.k(19); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 6), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j16_20 += 1.0 * d; }
    { const double k = m.k(20); const double ya = fmax((double)m.at(yoff + 16), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j16_16 += -1.0 * d; }
    { const double k = m.k(20); const double ya = fmax((double)m.at(yoff + 16), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j16_26 += -1.0 * d; }
    { const double k = m.k(72); const double ya = fmax((double)m.at(yoff + 16), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j16_16 += -1.0 * d; }
    { const double k = m.k(72); const double ya = fmax((double)m.at(yoff + 16), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j16_24 += -1.0 * d; }
    { const double k = m.k(14); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j17_22 += 1.0 * d; }
    { const double k = m.k(14); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j17_26 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(15); const double ya = fmax((double)m.at(yoff + 11), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j17_17 += -1.0 * d; }
    { const double k = m.k(15); const double ya = fmax((double)m.at(yoff + 11), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j17_26 += -1.0 * d; }
    { const double k = m.k(16); const double ya = fmax((double)m.at(yoff + 11), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j17_17 += -1.0 * d; }
    { const double k = m.k(16); const double ya = fmax((double)m.at(yoff + 11), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j17_24 += -1.0 * d; }
    { const double k = m.k(35); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 8), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j17_20 += 1.0 * d; }
    { const double k = m.k(35); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 8), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j17_27 += 1.0 * d; }
    m.J(60) = j16_16;
    m.J(61) = j16_20;
    m.J(62) = j16_24;
    m.J(63) = j16_26;
    m.J(64) = j16_29;
    m.J(65) = j17_17;
    m.J(66) = j17_20;
    m.J(67) = j17_22;
    m.J(68) = j17_24;
    m.J(69) = j17_26;
    m.J(70) = j17_27;
  }
  P3D_FENCE();
  {  // rows HCO+, CH
    double j18_18 = 0.0;
    double j18_19 = 0.0;
    double j18_20 = 0.0;
    double j18_21 = 0.0;
    double j18_22 = 0.0;
    double j18_24 = 0.0;
    double j18_26 = 0.0;
    double j18_28 = 0.0;
    double j18_29 = 0.0;
    double j18_30 = 0.0;
    double j19_11 = 0.0;
    double j19_19 = 0.0;
    double j19_24 = 0.0;
    double j19_25 = 0.0;
    double j19_29 = 0.0;
    { const double k = m.k(11); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j18_22 += 1.0 * d; }
    { const double k = m.k(11); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j18_21 += 1.0 * d; }
    { const double k = m.k(29); const double ya = fmax((double)m.at(yoff + 9), 1e-30); const double yb = fmax((double)m.at(yoff + 6), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j18_28 += 1.0 * d; }
    { const double k = m.k(29); const double ya = fmax((double)m.at(yoff + 9), 1e-30); const double yb = fmax((double)m.at(yoff + 6), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j18_20 += 1.0 * d; }
    { const double k = m.k(30); const double ya = fmax((double)m.at(yoff + 22), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j18_18 += -1.0 * d; }
    { const double k = m.k(30); const double ya = fmax((double)m.at(yoff + 22), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j18_24 += -1.0 * d; }
    { const double k = m.k(31); const double ya = fmax((double)m.at(yoff + 22), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j18_18 += -1.0 * d; }
    { const double k = m.k(43); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j18_30 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(43); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j18_26 += 1.0 * d; }
    { const double k = m.k(65); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 13), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j18_19 += 1.0 * d; }
    { const double k = m.k(65); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 13), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j18_29 += 1.0 * d; }
    { const double k = m.k(13); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 13), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j19_19 += -1.0 * d; }
    { const double k = m.k(13); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 13), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j19_29 += -1.0 * d; }
    { const double k = m.k(61); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j19_19 += -1.0 * d; }
    { const double k = m.k(61); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j19_25 += -1.0 * d; }
    { const double k = m.k(62); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j19_19 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(63); const double ya = fmax((double)m.at(yoff + 12), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j19_11 += 1.0 * d; }
    { const double k = m.k(63); const double ya = fmax((double)m.at(yoff + 12), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j19_24 += 1.0 * d; }
    { const double k = m.k(65); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 13), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j19_19 += -1.0 * d; }
    { const double k = m.k(65); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 13), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j19_29 += -1.0 * d; }
    m.J(71) = j18_18;
    m.J(72) = j18_19;
    m.J(73) = j18_20;
    m.J(74) = j18_21;
    m.J(75) = j18_22;
    m.J(76) = j18_24;
    m.J(77) = j18_26;
    m.J(78) = j18_28;
    m.J(79) = j18_29;
    m.J(80) = j18_30;
    m.J(81) = j19_11;
    m.J(82) = j19_19;
    m.J(83) = j19_24;
    m.J(84) = j19_25;
    m.J(85) = j19_29;
  }
  P3D_FENCE();
  {  // rows H3+, H2O
    double j20_0 = 0.0;
    double j20_20 = 0.0;
    double j20_24 = 0.0;
    double j20_26 = 0.0;
    double j20_27 = 0.0;
    double j20_28 = 0.0;
    double j20_29 = 0.0;
    double j21_12 = 0.0;
    double j21_21 = 0.0;
    double j21_22 = 0.0;
    double j21_23 = 0.0;
    double j21_24 = 0.0;
    double j21_26 = 0.0;
    double j21_27 = 0.0;
    { const double k = m.k(19); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 6), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j20_29 += -1.0 * d; }
    { const double k = m.k(19); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 6), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j20_20 += -1.0 * d; }
    { const double k = m.k(29); const double ya = fmax((double)m.at(yoff + 9), 1e-30); const double yb = fmax((double)m.at(yoff + 6), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j20_28 += -1.0 * d; }
    { const double k = m.k(29); const double ya = fmax((double)m.at(yoff + 9), 1e-30); const double yb = fmax((double)m.at(yoff + 6), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j20_20 += -1.0 * d; }
    { const double k = m.k(32); const double ya = fmax((double)m.at(yoff + 5), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j20_0 += 1.0 * d; }
    { const double k = m.k(32); const double ya = fmax((double)m.at(yoff + 5), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j20_26 += 1.0 * d; }
    { const double k = m.k(33); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j20_20 += -1.0 * d; }
    { const double k = m.k(33); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j20_24 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(34); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j20_20 += -1.0 * d; }
    { const double k = m.k(34); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j20_24 += -1.0 * d; }
    { const double k = m.k(35); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 8), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j20_20 += -1.0 * d; }
    { const double k = m.k(35); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 8), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j20_27 += -1.0 * d; }
    { const double k = m.k(11); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j21_22 += -1.0 * d; }
    { const double k = m.k(11); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j21_21 += -1.0 * d; }
    { const double k = m.k(23); const double ya = fmax((double)m.at(yoff + 19), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j21_12 += 1.0 * d; }
    { const double k = m.k(23); const double ya = fmax((double)m.at(yoff + 19), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j21_24 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(25); const double ya = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j21_21 += -1.0 * d; }
    { const double k = m.k(28); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j21_23 += 1.0 * d; }
    { const double k = m.k(28); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j21_26 += 1.0 * d; }
    { const double k = m.k(53); const double ya = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j21_21 += -1.0 * d; }
    { const double k = m.k(67); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j21_27 += -1.0 * d; }
    { const double k = m.k(67); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j21_21 += -1.0 * d; }
    m.J(86) = j20_0;
    m.J(87) = j20_20;
    m.J(88) = j20_24;
    m.J(89) = j20_26;
    m.J(90) = j20_27;
    m.J(91) = j20_28;
    m.J(92) = j20_29;
    m.J(93) = j21_12;
    m.J(94) = j21_21;
    m.J(95) = j21_22;
    m.J(96) = j21_23;
    m.J(97) = j21_24;
    m.J(98) = j21_26;
    m.J(99) = j21_27;
  }
  P3D_FENCE();
  {  // rows C+, OH
    double j22_9 = 0.0;
    double j22_15 = 0.0;
    double j22_21 = 0.0;
    double j22_22 = 0.0;
    double j22_23 = 0.0;
    double j22_24 = 0.0;
    double j22_26 = 0.0;
    double j22_27 = 0.0;
    double j22_28 = 0.0;
    double j23_12 = 0.0;
    double j23_14 = 0.0;
    double j23_21 = 0.0;
    double j23_22 = 0.0;
    double j23_23 = 0.0;
    double j23_24 = 0.0;
    double j23_25 = 0.0;
    double j23_26 = 0.0;
    double j23_27 = 0.0;
    double j23_29 = 0.0;
    { const double k = m.k(7); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j22_27 += 1.0 * d; }
    { const double k = m.k(8); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j22_22 += -1.0 * d; }
    { const double k = m.k(8); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j22_24 += -1.0 * d; }
    { const double k = m.k(10); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j22_22 += -1.0 * d; }
    { const double k = m.k(10); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j22_23 += -1.0 * d; }
    { const double k = m.k(11); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j22_22 += -1.0 * d; }
    { const double k = m.k(11); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j22_21 += -1.0 * d; }
    { const double k = m.k(14); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j22_22 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(14); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j22_26 += -1.0 * d; }
    { const double k = m.k(45); const double ya = fmax((double)m.at(yoff + 30), 1e-30); const double yb = fmax((double)m.at(yoff + 9), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j22_9 += 1.0 * d; }
    { const double k = m.k(45); const double ya = fmax((double)m.at(yoff + 30), 1e-30); const double yb = fmax((double)m.at(yoff + 9), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j22_28 += 1.0 * d; }
    { const double k = m.k(47); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double d = k * pm * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j22_22 += -1.0 * d; }
    { const double k = m.k(47); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double d = k * ya * p.nH * dpm * (ye > 1e-30 ? 1.0 : 0.0); j22_24 += -1.0 * d; }
    { const double k = m.k(55); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j22_27 += 1.0 * d; }
    { const double k = m.k(60); const double ya = fmax((double)m.at(yoff + 20), 1e-30); const double yb = fmax((double)m.at(yoff + 7), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j22_15 += -1.0 * d; }
    { const double k = m.k(60); const double ya = fmax((double)m.at(yoff + 20), 1e-30); const double yb = fmax((double)m.at(yoff + 7), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j22_22 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(10); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j23_22 += -1.0 * d; }
    { const double k = m.k(10); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j23_23 += -1.0 * d; }
    { const double k = m.k(12); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j23_27 += -1.0 * d; }
    { const double k = m.k(12); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j23_23 += -1.0 * d; }
    { const double k = m.k(22); const double ya = fmax((double)m.at(yoff + 19), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j23_12 += 1.0 * d; }
    { const double k = m.k(22); const double ya = fmax((double)m.at(yoff + 19), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j23_24 += 1.0 * d; }
    { const double k = m.k(24); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j23_23 += -1.0 * d; }
    { const double k = m.k(25); const double ya = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j23_21 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(26); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j23_29 += 1.0 * d; }
    { const double k = m.k(26); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j23_26 += 1.0 * d; }
    { const double k = m.k(27); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j23_23 += -1.0 * d; }
    { const double k = m.k(27); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j23_25 += -1.0 * d; }
    { const double k = m.k(28); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j23_23 += -1.0 * d; }
    { const double k = m.k(28); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j23_26 += -1.0 * d; }
    { const double k = m.k(53); const double ya = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j23_21 += 1.0 * d; }
    { const double k = m.k(54); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j23_23 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(56); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j23_29 += -1.0 * d; }
    { const double k = m.k(56); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j23_23 += -1.0 * d; }
    { const double k = m.k(74); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j23_14 += 1.0 * d; }
    { const double k = m.k(74); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j23_24 += 1.0 * d; }
    m.J(100) = j22_9;
    m.J(101) = j22_15;
    m.J(102) = j22_21;
    m.J(103) = j22_22;
    m.J(104) = j22_23;
    m.J(105) = j22_24;
    m.J(106) = j22_26;
    m.J(107) = j22_27;
    m.J(108) = j22_28;
    m.J(109) = j23_12;
    m.J(110) = j23_14;
    m.J(111) = j23_21;
    m.J(112) = j23_22;
    m.J(113) = j23_23;
    m.J(114) = j23_24;
    m.J(115) = j23_25;
    m.J(116) = j23_26;
    m.J(117) = j23_27;
    m.J(118) = j23_29;
  }
  P3D_FENCE();
  {  // rows e-, H
    double j24_1 = 0.0;
    double j24_2 = 0.0;
    double j24_3 = 0.0;
    double j24_4 = 0.0;
    double j24_5 = 0.0;
    double j24_6 = 0.0;
    double j24_7 = 0.0;
    double j24_8 = 0.0;
    double j24_9 = 0.0;
    double j24_11 = 0.0;
    double j24_12 = 0.0;
    double j24_13 = 0.0;
    double j24_14 = 0.0;
    double j24_16 = 0.0;
    double j24_17 = 0.0;
    double j24_18 = 0.0;
    double j24_19 = 0.0;
    double j24_20 = 0.0;
    double j24_22 = 0.0;
    double j24_24 = 0.0;
    double j24_25 = 0.0;
    double j24_26 = 0.0;
    double j24_27 = 0.0;
    double j24_29 = 0.0;
    double j24_30 = 0.0;
    double j25_0 = 0.0;
    double j25_7 = 0.0;
    double j25_10 = 0.0;
    double j25_11 = 0.0;
    double j25_12 = 0.0;
    double j25_13 = 0.0;
    double j25_14 = 0.0;
    double j25_16 = 0.0;
    double j25_17 = 0.0;
    double j25_18 = 0.0;
    double j25_19 = 0.0;
    double j25_20 = 0.0;
    double j25_21 = 0.0;
    double j25_22 = 0.0;
    double j25_23 = 0.0;
    double j25_24 = 0.0;
    double j25_25 = 0.0;
    double j25_26 = 0.0;
    double j25_27 = 0.0;
    double j25_29 = 0.0;
    double j25_30 = 0.0;
    { const double k = m.k(2); const double ya = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j24_26 += 1.0 * d; }
    { const double k = m.k(3); const double ya = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j24_25 += 1.0 * d; }
    { const double k = m.k(4); const double ya = fmax((double)m.at(yoff + 3), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_13 += -1.0 * d; }
    { const double k = m.k(4); const double ya = fmax((double)m.at(yoff + 3), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(5); const double ya = fmax((double)m.at(yoff + 1), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_25 += -1.0 * d; }
    { const double k = m.k(5); const double ya = fmax((double)m.at(yoff + 1), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(6); const double ya = fmax((double)m.at(yoff + 4), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_7 += 1.0 * d; }
    { const double k = m.k(6); const double ya = fmax((double)m.at(yoff + 4), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_25 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(7); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j24_27 += 1.0 * d; }
    { const double k = m.k(8); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_22 += -1.0 * d; }
    { const double k = m.k(8); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(16); const double ya = fmax((double)m.at(yoff + 11), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_17 += -1.0 * d; }
    { const double k = m.k(16); const double ya = fmax((double)m.at(yoff + 11), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(22); const double ya = fmax((double)m.at(yoff + 19), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_12 += -1.0 * d; }
    { const double k = m.k(22); const double ya = fmax((double)m.at(yoff + 19), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(23); const double ya = fmax((double)m.at(yoff + 19), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_12 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(23); const double ya = fmax((double)m.at(yoff + 19), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(30); const double ya = fmax((double)m.at(yoff + 22), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_18 += -1.0 * d; }
    { const double k = m.k(30); const double ya = fmax((double)m.at(yoff + 22), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(33); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_20 += -1.0 * d; }
    { const double k = m.k(33); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(34); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_20 += -1.0 * d; }
    { const double k = m.k(34); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(36); const double ya = fmax((double)m.at(yoff + 23), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j24_1 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(37); const double ya = fmax((double)m.at(yoff + 24), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_2 += -1.0 * d; }
    { const double k = m.k(37); const double ya = fmax((double)m.at(yoff + 24), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(38); const double ya = fmax((double)m.at(yoff + 25), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j24_3 += 1.0 * d; }
    { const double k = m.k(39); const double ya = fmax((double)m.at(yoff + 26), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_4 += -1.0 * d; }
    { const double k = m.k(39); const double ya = fmax((double)m.at(yoff + 26), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(40); const double ya = fmax((double)m.at(yoff + 27), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j24_5 += 1.0 * d; }
    { const double k = m.k(41); const double ya = fmax((double)m.at(yoff + 28), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_6 += -1.0 * d; }
    { const double k = m.k(41); const double ya = fmax((double)m.at(yoff + 28), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(44); const double ya = fmax((double)m.at(yoff + 29), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j24_8 += 1.0 * d; }
    { const double k = m.k(46); const double ya = fmax((double)m.at(yoff + 30), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_9 += -1.0 * d; }
    { const double k = m.k(46); const double ya = fmax((double)m.at(yoff + 30), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(55); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j24_27 += 1.0 * d; }
    { const double k = m.k(63); const double ya = fmax((double)m.at(yoff + 12), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_11 += -1.0 * d; }
    { const double k = m.k(63); const double ya = fmax((double)m.at(yoff + 12), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(64); const double ya = fmax((double)m.at(yoff + 12), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_11 += -1.0 * d; }
    { const double k = m.k(64); const double ya = fmax((double)m.at(yoff + 12), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(65); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 13), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_19 += 1.0 * d; }
    { const double k = m.k(65); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 13), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_29 += 1.0 * d; }
    { const double k = m.k(66); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_30 += -1.0 * d; }
    { const double k = m.k(66); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(70); const double ya = fmax((double)m.at(yoff + 28), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_6 += -1.0 * d; }
    { const double k = m.k(70); const double ya = fmax((double)m.at(yoff + 28), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(71); const double ya = fmax((double)m.at(yoff + 26), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_4 += -1.0 * d; }
    { const double k = m.k(71); const double ya = fmax((double)m.at(yoff + 26), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(72); const double ya = fmax((double)m.at(yoff + 16), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_16 += -1.0 * d; }
    { const double k = m.k(72); const double ya = fmax((double)m.at(yoff + 16), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(73); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_14 += -1.0 * d; }
    { const double k = m.k(73); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(74); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j24_14 += -1.0 * d; }
    { const double k = m.k(74); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j24_24 += -1.0 * d; }
    { const double k = m.k(0); const double ya = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * p.nH * 0.5 * (ya > 1e-30 ? 1.0 : 0.0); j25_25 += -2.0 * d; }
    { const double k = m.k(1); const double ya = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j25_26 += 2.0 * d; }
    P3D_FENCE();
    { const double k = m.k(3); const double ya = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j25_25 += -1.0 * d; }
    { const double k = m.k(4); const double ya = fmax((double)m.at(yoff + 3), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_13 += 1.0 * d; }
    { const double k = m.k(4); const double ya = fmax((double)m.at(yoff + 3), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_24 += 1.0 * d; }
    { const double k = m.k(5); const double ya = fmax((double)m.at(yoff + 1), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_25 += -1.0 * d; }
    { const double k = m.k(5); const double ya = fmax((double)m.at(yoff + 1), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_24 += -1.0 * d; }
    { const double k = m.k(6); const double ya = fmax((double)m.at(yoff + 4), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_7 += -1.0 * d; }
    { const double k = m.k(6); const double ya = fmax((double)m.at(yoff + 4), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_25 += -1.0 * d; }
    { const double k = m.k(11); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_22 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(11); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_21 += 1.0 * d; }
    { const double k = m.k(12); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_27 += 1.0 * d; }
    { const double k = m.k(12); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_23 += 1.0 * d; }
    { const double k = m.k(13); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 13), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_19 += 1.0 * d; }
    { const double k = m.k(13); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 13), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_29 += 1.0 * d; }
    { const double k = m.k(14); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_22 += 1.0 * d; }
    { const double k = m.k(14); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_26 += 1.0 * d; }
    { const double k = m.k(15); const double ya = fmax((double)m.at(yoff + 11), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_17 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(15); const double ya = fmax((double)m.at(yoff + 11), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_26 += 1.0 * d; }
    { const double k = m.k(16); const double ya = fmax((double)m.at(yoff + 11), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_17 += 1.0 * d; }
    { const double k = m.k(16); const double ya = fmax((double)m.at(yoff + 11), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_24 += 1.0 * d; }
    { const double k = m.k(17); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 3), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_29 += 1.0 * d; }
    { const double k = m.k(17); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 3), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_13 += 1.0 * d; }
    { const double k = m.k(18); const double ya = fmax((double)m.at(yoff + 14), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_10 += -1.0 * d; }
    { const double k = m.k(18); const double ya = fmax((double)m.at(yoff + 14), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_25 += -1.0 * d; }
    { const double k = m.k(20); const double ya = fmax((double)m.at(yoff + 16), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_16 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(20); const double ya = fmax((double)m.at(yoff + 16), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_26 += 1.0 * d; }
    { const double k = m.k(21); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_14 += 1.0 * d; }
    { const double k = m.k(21); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_26 += 1.0 * d; }
    { const double k = m.k(22); const double ya = fmax((double)m.at(yoff + 19), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_12 += 2.0 * d; }
    { const double k = m.k(22); const double ya = fmax((double)m.at(yoff + 19), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_24 += 2.0 * d; }
    { const double k = m.k(23); const double ya = fmax((double)m.at(yoff + 19), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_12 += 1.0 * d; }
    { const double k = m.k(23); const double ya = fmax((double)m.at(yoff + 19), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_24 += 1.0 * d; }
    { const double k = m.k(24); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j25_23 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(25); const double ya = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j25_21 += 1.0 * d; }
    { const double k = m.k(26); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_29 += 1.0 * d; }
    { const double k = m.k(26); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_26 += 1.0 * d; }
    { const double k = m.k(27); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_23 += -1.0 * d; }
    { const double k = m.k(27); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_25 += -1.0 * d; }
    { const double k = m.k(28); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_23 += 1.0 * d; }
    { const double k = m.k(28); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_26 += 1.0 * d; }
    { const double k = m.k(30); const double ya = fmax((double)m.at(yoff + 22), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_18 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(30); const double ya = fmax((double)m.at(yoff + 22), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_24 += 1.0 * d; }
    { const double k = m.k(31); const double ya = fmax((double)m.at(yoff + 22), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j25_18 += 1.0 * d; }
    { const double k = m.k(32); const double ya = fmax((double)m.at(yoff + 5), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_0 += 1.0 * d; }
    { const double k = m.k(32); const double ya = fmax((double)m.at(yoff + 5), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_26 += 1.0 * d; }
    { const double k = m.k(33); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_20 += 1.0 * d; }
    { const double k = m.k(33); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_24 += 1.0 * d; }
    { const double k = m.k(34); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_20 += 3.0 * d; }
    { const double k = m.k(34); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_24 += 3.0 * d; }
    P3D_FENCE();
    { const double k = m.k(42); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_30 += -1.0 * d; }
    { const double k = m.k(42); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_25 += -1.0 * d; }
    { const double k = m.k(43); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_30 += 1.0 * d; }
    { const double k = m.k(43); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_26 += 1.0 * d; }
    { const double k = m.k(49); const double ya = fmax((double)m.at(yoff + 3), 1e-30); const double d = k * pm * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_13 += 1.0 * d; }
    { const double k = m.k(49); const double ya = fmax((double)m.at(yoff + 3), 1e-30); const double d = k * ya * p.nH * dpm * (ye > 1e-30 ? 1.0 : 0.0); j25_24 += 1.0 * d; }
    { const double k = m.k(53); const double ya = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j25_21 += 1.0 * d; }
    { const double k = m.k(54); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j25_23 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(56); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_29 += 1.0 * d; }
    { const double k = m.k(56); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_23 += 1.0 * d; }
    { const double k = m.k(61); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_19 += -1.0 * d; }
    { const double k = m.k(61); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_25 += -1.0 * d; }
    { const double k = m.k(62); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j25_19 += 1.0 * d; }
    { const double k = m.k(63); const double ya = fmax((double)m.at(yoff + 12), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_11 += 1.0 * d; }
    { const double k = m.k(63); const double ya = fmax((double)m.at(yoff + 12), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_24 += 1.0 * d; }
    { const double k = m.k(64); const double ya = fmax((double)m.at(yoff + 12), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_11 += 2.0 * d; }
    P3D_FENCE();
    { const double k = m.k(64); const double ya = fmax((double)m.at(yoff + 12), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_24 += 2.0 * d; }
    { const double k = m.k(67); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_27 += 2.0 * d; }
    { const double k = m.k(67); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_21 += 2.0 * d; }
    { const double k = m.k(68); const double ya = fmax((double)m.at(yoff + 2), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_26 += 2.0 * d; }
    { const double k = m.k(68); const double ya = fmax((double)m.at(yoff + 2), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_25 += 2.0 * d; }
    { const double k = m.k(69); const double ya = fmax((double)m.at(yoff + 2), 1e-30); const double d = 2.0 * k * ya * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_26 += 2.0 * d; }
    { const double k = m.k(72); const double ya = fmax((double)m.at(yoff + 16), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_16 += 1.0 * d; }
    { const double k = m.k(72); const double ya = fmax((double)m.at(yoff + 16), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_24 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(73); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_14 += 2.0 * d; }
    { const double k = m.k(73); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_24 += 2.0 * d; }
    { const double k = m.k(74); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j25_14 += 1.0 * d; }
    { const double k = m.k(74); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j25_24 += 1.0 * d; }
    m.J(119) = j24_1;
    m.J(120) = j24_2;
    m.J(121) = j24_3;
    m.J(122) = j24_4;
    m.J(123) = j24_5;
    m.J(124) = j24_6;
    m.J(125) = j24_7;
    m.J(126) = j24_8;
    m.J(127) = j24_9;
    m.J(128) = j24_11;
    m.J(129) = j24_12;
    m.J(130) = j24_13;
    m.J(131) = j24_14;
    m.J(132) = j24_16;
    m.J(133) = j24_17;
    m.J(134) = j24_18;
    m.J(135) = j24_19;
    m.J(136) = j24_20;
    m.J(137) = j24_22;
    m.J(138) = j24_24;
    m.J(139) = j24_25;
    m.J(140) = j24_26;
    m.J(141) = j24_27;
    m.J(142) = j24_29;
    m.J(143) = j24_30;
    m.J(144) = j25_0;
    m.J(145) = j25_7;
    m.J(146) = j25_10;
    m.J(147) = j25_11;
    m.J(148) = j25_12;
    m.J(149) = j25_13;
    m.J(150) = j25_14;
    m.J(151) = j25_16;
    m.J(152) = j25_17;
    m.J(153) = j25_18;
    m.J(154) = j25_19;
    m.J(155) = j25_20;
    m.J(156) = j25_21;
    m.J(157) = j25_22;
    m.J(158) = j25_23;
    m.J(159) = j25_24;
    m.J(160) = j25_25;
    m.J(161) = j25_26;
    m.J(162) = j25_27;
    m.J(163) = j25_29;
    m.J(164) = j25_30;
  }
  P3D_FENCE();
  {  // rows H2, C
    double j26_0 = 0.0;
    double j26_7 = 0.0;
    double j26_14 = 0.0;
    double j26_16 = 0.0;
    double j26_17 = 0.0;
    double j26_19 = 0.0;
    double j26_20 = 0.0;
    double j26_22 = 0.0;
    double j26_23 = 0.0;
    double j26_24 = 0.0;
    double j26_25 = 0.0;
    double j26_26 = 0.0;
    double j26_27 = 0.0;
    double j26_28 = 0.0;
    double j26_29 = 0.0;
    double j26_30 = 0.0;
    double j27_11 = 0.0;
    double j27_15 = 0.0;
    double j27_17 = 0.0;
    double j27_19 = 0.0;
    double j27_20 = 0.0;
    double j27_21 = 0.0;
    double j27_22 = 0.0;
    double j27_23 = 0.0;
    double j27_24 = 0.0;
    double j27_25 = 0.0;
    double j27_27 = 0.0;
    double j27_28 = 0.0;
    double j27_30 = 0.0;
    { const double k = m.k(0); const double ya = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * p.nH * 0.5 * (ya > 1e-30 ? 1.0 : 0.0); j26_25 += 1.0 * d; }
    { const double k = m.k(1); const double ya = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j26_26 += -1.0 * d; }
    { const double k = m.k(2); const double ya = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j26_26 += -1.0 * d; }
    { const double k = m.k(6); const double ya = fmax((double)m.at(yoff + 4), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j26_7 += 1.0 * d; }
    { const double k = m.k(6); const double ya = fmax((double)m.at(yoff + 4), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j26_25 += 1.0 * d; }
    { const double k = m.k(14); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j26_22 += -1.0 * d; }
    { const double k = m.k(14); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j26_26 += -1.0 * d; }
    { const double k = m.k(15); const double ya = fmax((double)m.at(yoff + 11), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j26_17 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(15); const double ya = fmax((double)m.at(yoff + 11), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j26_26 += -1.0 * d; }
    { const double k = m.k(19); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 6), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j26_29 += 1.0 * d; }
    { const double k = m.k(19); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 6), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j26_20 += 1.0 * d; }
    { const double k = m.k(20); const double ya = fmax((double)m.at(yoff + 16), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j26_16 += -1.0 * d; }
    { const double k = m.k(20); const double ya = fmax((double)m.at(yoff + 16), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j26_26 += -1.0 * d; }
    { const double k = m.k(21); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j26_14 += -1.0 * d; }
    { const double k = m.k(21); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j26_26 += -1.0 * d; }
    { const double k = m.k(26); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j26_29 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(26); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j26_26 += -1.0 * d; }
    { const double k = m.k(27); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j26_23 += 1.0 * d; }
    { const double k = m.k(27); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j26_25 += 1.0 * d; }
    { const double k = m.k(28); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j26_23 += -1.0 * d; }
    { const double k = m.k(28); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j26_26 += -1.0 * d; }
    { const double k = m.k(29); const double ya = fmax((double)m.at(yoff + 9), 1e-30); const double yb = fmax((double)m.at(yoff + 6), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j26_28 += 1.0 * d; }
    { const double k = m.k(29); const double ya = fmax((double)m.at(yoff + 9), 1e-30); const double yb = fmax((double)m.at(yoff + 6), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j26_20 += 1.0 * d; }
    { const double k = m.k(32); const double ya = fmax((double)m.at(yoff + 5), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j26_0 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(32); const double ya = fmax((double)m.at(yoff + 5), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j26_26 += -1.0 * d; }
    { const double k = m.k(33); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j26_20 += 1.0 * d; }
    { const double k = m.k(33); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j26_24 += 1.0 * d; }
    { const double k = m.k(35); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 8), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j26_20 += 1.0 * d; }
    { const double k = m.k(35); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 8), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j26_27 += 1.0 * d; }
    { const double k = m.k(43); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j26_30 += -1.0 * d; }
    { const double k = m.k(43); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j26_26 += -1.0 * d; }
    { const double k = m.k(61); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j26_19 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(61); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j26_25 += 1.0 * d; }
    { const double k = m.k(68); const double ya = fmax((double)m.at(yoff + 2), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j26_26 += -1.0 * d; }
    { const double k = m.k(68); const double ya = fmax((double)m.at(yoff + 2), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j26_25 += -1.0 * d; }
    { const double k = m.k(69); const double ya = fmax((double)m.at(yoff + 2), 1e-30); const double d = 2.0 * k * ya * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j26_26 += -1.0 * d; }
    { const double k = m.k(7); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j27_27 += -1.0 * d; }
    { const double k = m.k(8); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j27_22 += 1.0 * d; }
    { const double k = m.k(8); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j27_24 += 1.0 * d; }
    { const double k = m.k(9); const double ya = fmax((double)m.at(yoff + 9), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j27_28 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(12); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j27_27 += -1.0 * d; }
    { const double k = m.k(12); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j27_23 += -1.0 * d; }
    { const double k = m.k(16); const double ya = fmax((double)m.at(yoff + 11), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j27_17 += 1.0 * d; }
    { const double k = m.k(16); const double ya = fmax((double)m.at(yoff + 11), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j27_24 += 1.0 * d; }
    { const double k = m.k(35); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 8), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j27_20 += -1.0 * d; }
    { const double k = m.k(35); const double ya = fmax((double)m.at(yoff + 6), 1e-30); const double yb = fmax((double)m.at(yoff + 8), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j27_27 += -1.0 * d; }
    { const double k = m.k(47); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double d = k * pm * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j27_22 += 1.0 * d; }
    { const double k = m.k(47); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double d = k * ya * p.nH * dpm * (ye > 1e-30 ? 1.0 : 0.0); j27_24 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(52); const double ya = fmax((double)m.at(yoff + 9), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j27_28 += 1.0 * d; }
    { const double k = m.k(55); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j27_27 += -1.0 * d; }
    { const double k = m.k(57); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 20), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j27_27 += -1.0 * d; }
    { const double k = m.k(57); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 20), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j27_15 += -1.0 * d; }
    { const double k = m.k(61); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j27_19 += 1.0 * d; }
    { const double k = m.k(61); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j27_25 += 1.0 * d; }
    { const double k = m.k(62); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j27_19 += 1.0 * d; }
    { const double k = m.k(64); const double ya = fmax((double)m.at(yoff + 12), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j27_11 += 1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(64); const double ya = fmax((double)m.at(yoff + 12), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j27_24 += 1.0 * d; }
    { const double k = m.k(66); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j27_30 += 1.0 * d; }
    { const double k = m.k(66); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j27_24 += 1.0 * d; }
    { const double k = m.k(67); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j27_27 += -1.0 * d; }
    { const double k = m.k(67); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j27_21 += -1.0 * d; }
    m.J(165) = j26_0;
    m.J(166) = j26_7;
    m.J(167) = j26_14;
    m.J(168) = j26_16;
    m.J(169) = j26_17;
    m.J(170) = j26_19;
    m.J(171) = j26_20;
    m.J(172) = j26_22;
    m.J(173) = j26_23;
    m.J(174) = j26_24;
    m.J(175) = j26_25;
    m.J(176) = j26_26;
    m.J(177) = j26_27;
    m.J(178) = j26_28;
    m.J(179) = j26_29;
    m.J(180) = j26_30;
    m.J(181) = j27_11;
    m.J(182) = j27_15;
    m.J(183) = j27_17;
    m.J(184) = j27_19;
    m.J(185) = j27_20;
    m.J(186) = j27_21;
    m.J(187) = j27_22;
    m.J(188) = j27_23;
    m.J(189) = j27_24;
    m.J(190) = j27_25;
    m.J(191) = j27_27;
    m.J(192) = j27_28;
    m.J(193) = j27_30;
  }
  P3D_FENCE();
  {  // rows CO, O
    double j28_9 = 0.0;
    double j28_15 = 0.0;
    double j28_18 = 0.0;
    double j28_19 = 0.0;
    double j28_20 = 0.0;
    double j28_21 = 0.0;
    double j28_22 = 0.0;
    double j28_23 = 0.0;
    double j28_24 = 0.0;
    double j28_25 = 0.0;
    double j28_27 = 0.0;
    double j28_28 = 0.0;
    double j28_29 = 0.0;
    double j28_30 = 0.0;
    double j29_9 = 0.0;
    double j29_10 = 0.0;
    double j29_13 = 0.0;
    double j29_14 = 0.0;
    double j29_15 = 0.0;
    double j29_16 = 0.0;
    double j29_19 = 0.0;
    double j29_20 = 0.0;
    double j29_22 = 0.0;
    double j29_23 = 0.0;
    double j29_24 = 0.0;
    double j29_25 = 0.0;
    double j29_26 = 0.0;
    double j29_27 = 0.0;
    double j29_28 = 0.0;
    double j29_29 = 0.0;
    double j29_30 = 0.0;
    { const double k = m.k(9); const double ya = fmax((double)m.at(yoff + 9), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j28_28 += -1.0 * d; }
    { const double k = m.k(10); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j28_22 += 1.0 * d; }
    { const double k = m.k(10); const double ya = fmax((double)m.at(yoff + 7), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j28_23 += 1.0 * d; }
    { const double k = m.k(12); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j28_27 += 1.0 * d; }
    { const double k = m.k(12); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j28_23 += 1.0 * d; }
    { const double k = m.k(13); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 13), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j28_19 += 1.0 * d; }
    { const double k = m.k(13); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 13), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j28_29 += 1.0 * d; }
    { const double k = m.k(29); const double ya = fmax((double)m.at(yoff + 9), 1e-30); const double yb = fmax((double)m.at(yoff + 6), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j28_28 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(29); const double ya = fmax((double)m.at(yoff + 9), 1e-30); const double yb = fmax((double)m.at(yoff + 6), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j28_20 += -1.0 * d; }
    { const double k = m.k(30); const double ya = fmax((double)m.at(yoff + 22), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j28_18 += 1.0 * d; }
    { const double k = m.k(30); const double ya = fmax((double)m.at(yoff + 22), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j28_24 += 1.0 * d; }
    { const double k = m.k(42); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j28_30 += 1.0 * d; }
    { const double k = m.k(42); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j28_25 += 1.0 * d; }
    { const double k = m.k(45); const double ya = fmax((double)m.at(yoff + 30), 1e-30); const double yb = fmax((double)m.at(yoff + 9), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j28_9 += -1.0 * d; }
    { const double k = m.k(45); const double ya = fmax((double)m.at(yoff + 30), 1e-30); const double yb = fmax((double)m.at(yoff + 9), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j28_28 += -1.0 * d; }
    { const double k = m.k(52); const double ya = fmax((double)m.at(yoff + 9), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j28_28 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(57); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 20), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j28_27 += 1.0 * d; }
    { const double k = m.k(57); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 20), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j28_15 += 1.0 * d; }
    { const double k = m.k(67); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j28_27 += 1.0 * d; }
    { const double k = m.k(67); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 17), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j28_21 += 1.0 * d; }
    { const double k = m.k(9); const double ya = fmax((double)m.at(yoff + 9), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j29_28 += 1.0 * d; }
    { const double k = m.k(13); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 13), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j29_19 += -1.0 * d; }
    { const double k = m.k(13); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 13), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j29_29 += -1.0 * d; }
    { const double k = m.k(17); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 3), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j29_29 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(17); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 3), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j29_13 += -1.0 * d; }
    { const double k = m.k(18); const double ya = fmax((double)m.at(yoff + 14), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j29_10 += 1.0 * d; }
    { const double k = m.k(18); const double ya = fmax((double)m.at(yoff + 14), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j29_25 += 1.0 * d; }
    { const double k = m.k(19); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 6), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j29_29 += -1.0 * d; }
    { const double k = m.k(19); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 6), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j29_20 += -1.0 * d; }
    { const double k = m.k(24); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j29_23 += 1.0 * d; }
    { const double k = m.k(26); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j29_29 += -1.0 * d; }
    { const double k = m.k(26); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j29_26 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(27); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j29_23 += 1.0 * d; }
    { const double k = m.k(27); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j29_25 += 1.0 * d; }
    { const double k = m.k(45); const double ya = fmax((double)m.at(yoff + 30), 1e-30); const double yb = fmax((double)m.at(yoff + 9), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j29_9 += 1.0 * d; }
    { const double k = m.k(45); const double ya = fmax((double)m.at(yoff + 30), 1e-30); const double yb = fmax((double)m.at(yoff + 9), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j29_28 += 1.0 * d; }
    { const double k = m.k(52); const double ya = fmax((double)m.at(yoff + 9), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j29_28 += 1.0 * d; }
    { const double k = m.k(54); const double ya = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j29_23 += 1.0 * d; }
    { const double k = m.k(56); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j29_29 += -1.0 * d; }
    { const double k = m.k(56); const double ya = fmax((double)m.at(yoff + 13), 1e-30); const double yb = fmax((double)m.at(yoff + 15), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j29_23 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(57); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 20), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j29_27 += 1.0 * d; }
    { const double k = m.k(57); const double ya = fmax((double)m.at(yoff + 8), 1e-30); const double yb = fmax((double)m.at(yoff + 20), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j29_15 += 1.0 * d; }
    { const double k = m.k(58); const double ya = fmax((double)m.at(yoff + 20), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j29_15 += 2.0 * d; }
    { const double k = m.k(59); const double ya = fmax((double)m.at(yoff + 20), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j29_15 += 2.0 * d; }
    { const double k = m.k(60); const double ya = fmax((double)m.at(yoff + 20), 1e-30); const double yb = fmax((double)m.at(yoff + 7), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j29_15 += 1.0 * d; }
    { const double k = m.k(60); const double ya = fmax((double)m.at(yoff + 20), 1e-30); const double yb = fmax((double)m.at(yoff + 7), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j29_22 += 1.0 * d; }
    { const double k = m.k(65); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 13), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j29_19 += -1.0 * d; }
    { const double k = m.k(65); const double ya = fmax((double)m.at(yoff + 10), 1e-30); const double yb = fmax((double)m.at(yoff + 13), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j29_29 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(66); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j29_30 += 1.0 * d; }
    { const double k = m.k(66); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j29_24 += 1.0 * d; }
    { const double k = m.k(72); const double ya = fmax((double)m.at(yoff + 16), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j29_16 += 1.0 * d; }
    { const double k = m.k(72); const double ya = fmax((double)m.at(yoff + 16), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j29_24 += 1.0 * d; }
    { const double k = m.k(73); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j29_14 += 1.0 * d; }
    { const double k = m.k(73); const double ya = fmax((double)m.at(yoff + 18), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j29_24 += 1.0 * d; }
    m.J(194) = j28_9;
    m.J(195) = j28_15;
    m.J(196) = j28_18;
    m.J(197) = j28_19;
    m.J(198) = j28_20;
    m.J(199) = j28_21;
    m.J(200) = j28_22;
    m.J(201) = j28_23;
    m.J(202) = j28_24;
    m.J(203) = j28_25;
    m.J(204) = j28_27;
    m.J(205) = j28_28;
    m.J(206) = j28_29;
    m.J(207) = j28_30;
    m.J(208) = j29_9;
    m.J(209) = j29_10;
    m.J(210) = j29_13;
    m.J(211) = j29_14;
    m.J(212) = j29_15;
    m.J(213) = j29_16;
    m.J(214) = j29_19;
    m.J(215) = j29_20;
    m.J(216) = j29_22;
    m.J(217) = j29_23;
    m.J(218) = j29_24;
    m.J(219) = j29_25;
    m.J(220) = j29_26;
    m.J(221) = j29_27;
    m.J(222) = j29_28;
    m.J(223) = j29_29;
    m.J(224) = j29_30;
  }
  P3D_FENCE();
  {  // rows CO+
    double j30_15 = 0.0;
    double j30_18 = 0.0;
    double j30_22 = 0.0;
    double j30_24 = 0.0;
    double j30_25 = 0.0;
    double j30_26 = 0.0;
    double j30_30 = 0.0;
    { const double k = m.k(31); const double ya = fmax((double)m.at(yoff + 22), 1e-30); const double d = k * (ya > 1e-30 ? 1.0 : 0.0); j30_18 += 1.0 * d; }
    { const double k = m.k(42); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j30_30 += -1.0 * d; }
    { const double k = m.k(42); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 1), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j30_25 += -1.0 * d; }
    { const double k = m.k(43); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j30_30 += -1.0 * d; }
    { const double k = m.k(43); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 2), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j30_26 += -1.0 * d; }
    { const double k = m.k(60); const double ya = fmax((double)m.at(yoff + 20), 1e-30); const double yb = fmax((double)m.at(yoff + 7), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j30_15 += 1.0 * d; }
    { const double k = m.k(60); const double ya = fmax((double)m.at(yoff + 20), 1e-30); const double yb = fmax((double)m.at(yoff + 7), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j30_22 += 1.0 * d; }
    { const double k = m.k(66); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * yb * p.nH * (ya > 1e-30 ? 1.0 : 0.0); j30_30 += -1.0 * d; }
    P3D_FENCE();
    { const double k = m.k(66); const double ya = fmax((double)m.at(yoff + 21), 1e-30); const double yb = fmax((double)m.at(yoff + 0), 1e-30); const double d = k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0); j30_24 += -1.0 * d; }
    m.J(225) = j30_15;
    m.J(226) = j30_18;
    m.J(227) = j30_22;
    m.J(228) = j30_24;
    m.J(229) = j30_25;
    m.J(230) = j30_26;
    m.J(231) = j30_30;
  }
  P3D_FENCE();
}

// M = I - c J factorised without pivoting in kPerm order (left-looking,
// row by row).  Stores the unit-lower multipliers, the upper factor and,
// on the diagonal, 1 / u_ii.
template <class Slot>
P3D_HD P3D_INLINE void lu_factor(const Slot &m, double c) {
  {  // level 1: H2+, S, Si, Fe, H-
    double a0_0 = 1.0 - c * m.J(0);
    double a0_26 = 0.0 - c * m.J(1);
    double a1_1 = 1.0 - c * m.J(2);
    double a1_2 = 0.0 - c * m.J(3);
    double a1_24 = 0.0 - c * m.J(4);
    double a3_3 = 1.0 - c * m.J(8);
    double a3_4 = 0.0 - c * m.J(9);
    double a3_24 = 0.0 - c * m.J(10);
    double a5_5 = 1.0 - c * m.J(14);
    double a5_6 = 0.0 - c * m.J(15);
    double a5_24 = 0.0 - c * m.J(16);
    double a7_7 = 1.0 - c * m.J(20);
    double a7_24 = 0.0 - c * m.J(21);
    double a7_25 = 0.0 - c * m.J(22);
    m.LU(0) = 1.0 / a0_0;
    m.LU(1) = a0_26;
    m.LU(2) = 1.0 / a1_1;
    m.LU(3) = a1_2;
    m.LU(4) = a1_24;
    m.LU(8) = 1.0 / a3_3;
    m.LU(9) = a3_4;
    m.LU(10) = a3_24;
    m.LU(14) = 1.0 / a5_5;
    m.LU(15) = a5_6;
    m.LU(16) = a5_24;
    m.LU(20) = 1.0 / a7_7;
    m.LU(21) = a7_24;
    m.LU(22) = a7_25;
  }
  P3D_FENCE();
  {  // level 1: He, O+, CH2+, H3O+, H2O+
    double a8_8 = 1.0 - c * m.J(23);
    double a8_9 = 0.0 - c * m.J(24);
    double a8_24 = 0.0 - c * m.J(25);
    double a8_28 = 0.0 - c * m.J(26);
    double a10_10 = 1.0 - c * m.J(31);
    double a10_13 = 0.0 - c * m.J(32);
    double a10_25 = 0.0 - c * m.J(33);
    double a10_29 = 0.0 - c * m.J(34);
    double a11_11 = 1.0 - c * m.J(35);
    double a11_17 = 0.0 - c * m.J(36);
    double a11_24 = 0.0 - c * m.J(37);
    double a11_26 = 0.0 - c * m.J(38);
    double a12_12 = 1.0 - c * m.J(39);
    double a12_14 = 0.0 - c * m.J(40);
    double a12_24 = 0.0 - c * m.J(41);
    double a12_26 = 0.0 - c * m.J(42);
    double a14_14 = 1.0 - c * m.J(51);
    double a14_16 = 0.0 - c * m.J(52);
    double a14_24 = 0.0 - c * m.J(53);
    double a14_26 = 0.0 - c * m.J(54);
    m.LU(23) = 1.0 / a8_8;
    m.LU(24) = a8_9;
    m.LU(25) = a8_24;
    m.LU(26) = a8_28;
    m.LU(31) = 1.0 / a10_10;
    m.LU(32) = a10_13;
    m.LU(33) = a10_25;
    m.LU(34) = a10_29;
    m.LU(35) = 1.0 / a11_11;
    m.LU(36) = a11_17;
    m.LU(37) = a11_24;
    m.LU(38) = a11_26;
    m.LU(39) = 1.0 / a12_12;
    m.LU(40) = a12_14;
    m.LU(41) = a12_24;
    m.LU(42) = a12_26;
    m.LU(51) = 1.0 / a14_14;
    m.LU(52) = a14_16;
    m.LU(53) = a14_24;
    m.LU(54) = a14_26;
  }
  P3D_FENCE();
  {  // level 1: O2, OH+, CH+, HCO+
    double a15_15 = 1.0 - c * m.J(55);
    double a15_22 = 0.0 - c * m.J(56);
    double a15_23 = 0.0 - c * m.J(57);
    double a15_27 = 0.0 - c * m.J(58);
    double a15_29 = 0.0 - c * m.J(59);
    double a16_16 = 1.0 - c * m.J(60);
    double a16_20 = 0.0 - c * m.J(61);
    double a16_24 = 0.0 - c * m.J(62);
    double a16_26 = 0.0 - c * m.J(63);
    double a16_29 = 0.0 - c * m.J(64);
    double a17_17 = 1.0 - c * m.J(65);
    double a17_20 = 0.0 - c * m.J(66);
    double a17_22 = 0.0 - c * m.J(67);
    double a17_24 = 0.0 - c * m.J(68);
    double a17_26 = 0.0 - c * m.J(69);
    double a17_27 = 0.0 - c * m.J(70);
    double a18_18 = 1.0 - c * m.J(71);
    double a18_19 = 0.0 - c * m.J(72);
    double a18_20 = 0.0 - c * m.J(73);
    double a18_21 = 0.0 - c * m.J(74);
    double a18_22 = 0.0 - c * m.J(75);
    double a18_24 = 0.0 - c * m.J(76);
    double a18_26 = 0.0 - c * m.J(77);
    double a18_28 = 0.0 - c * m.J(78);
    double a18_29 = 0.0 - c * m.J(79);
    double a18_30 = 0.0 - c * m.J(80);
    m.LU(55) = 1.0 / a15_15;
    m.LU(56) = a15_22;
    m.LU(57) = a15_23;
    m.LU(58) = a15_27;
    m.LU(59) = a15_29;
    m.LU(60) = 1.0 / a16_16;
    m.LU(61) = a16_20;
    m.LU(62) = a16_24;
    m.LU(63) = a16_26;
    m.LU(64) = a16_29;
    m.LU(65) = 1.0 / a17_17;
    m.LU(66) = a17_20;
    m.LU(67) = a17_22;
    m.LU(68) = a17_24;
    m.LU(69) = a17_26;
    m.LU(70) = a17_27;
    m.LU(71) = 1.0 / a18_18;
    m.LU(72) = a18_19;
    m.LU(73) = a18_20;
    m.LU(74) = a18_21;
    m.LU(75) = a18_22;
    m.LU(76) = a18_24;
    m.LU(77) = a18_26;
    m.LU(78) = a18_28;
    m.LU(79) = a18_29;
    m.LU(80) = a18_30;
  }
  P3D_FENCE();
  {  // level 2: S+, Si+, Fe+, He+, H+
    double a2_1 = 0.0 - c * m.J(5);
    double a2_2 = 1.0 - c * m.J(6);
    double a2_24 = 0.0 - c * m.J(7);
    double a4_3 = 0.0 - c * m.J(11);
    double a4_4 = 1.0 - c * m.J(12);
    double a4_24 = 0.0 - c * m.J(13);
    double a6_5 = 0.0 - c * m.J(17);
    double a6_6 = 1.0 - c * m.J(18);
    double a6_24 = 0.0 - c * m.J(19);
    double a9_8 = 0.0 - c * m.J(27);
    double a9_9 = 1.0 - c * m.J(28);
    double a9_24 = 0.0 - c * m.J(29);
    double a9_28 = 0.0 - c * m.J(30);
    double a13_10 = 0.0 - c * m.J(43);
    double a13_13 = 1.0 - c * m.J(44);
    double a13_22 = 0.0 - c * m.J(45);
    double a13_23 = 0.0 - c * m.J(46);
    double a13_24 = 0.0 - c * m.J(47);
    double a13_25 = 0.0 - c * m.J(48);
    double a13_29 = 0.0 - c * m.J(49);
    double a13_30 = 0.0 - c * m.J(50);
    a2_1 *= m.LU(2);
    a2_2 -= a2_1 * m.LU(3);
    a2_24 -= a2_1 * m.LU(4);
    a4_3 *= m.LU(8);
    a4_4 -= a4_3 * m.LU(9);
    a4_24 -= a4_3 * m.LU(10);
    a6_5 *= m.LU(14);
    a6_6 -= a6_5 * m.LU(15);
    a6_24 -= a6_5 * m.LU(16);
    a9_8 *= m.LU(23);
    a9_9 -= a9_8 * m.LU(24);
    a9_24 -= a9_8 * m.LU(25);
    a9_28 -= a9_8 * m.LU(26);
    a13_10 *= m.LU(31);
    a13_13 -= a13_10 * m.LU(32);
    a13_25 -= a13_10 * m.LU(33);
    a13_29 -= a13_10 * m.LU(34);
    m.LU(5) = a2_1;
    m.LU(6) = 1.0 / a2_2;
    m.LU(7) = a2_24;
    m.LU(11) = a4_3;
    m.LU(12) = 1.0 / a4_4;
    m.LU(13) = a4_24;
    m.LU(17) = a6_5;
    m.LU(18) = 1.0 / a6_6;
    m.LU(19) = a6_24;
    m.LU(27) = a9_8;
    m.LU(28) = 1.0 / a9_9;
    m.LU(29) = a9_24;
    m.LU(30) = a9_28;
    m.LU(43) = a13_10;
    m.LU(44) = 1.0 / a13_13;
    m.LU(45) = a13_22;
    m.LU(46) = a13_23;
    m.LU(47) = a13_24;
    m.LU(48) = a13_25;
    m.LU(49) = a13_29;
    m.LU(50) = a13_30;
  }
  P3D_FENCE();
  {  // level 2: CH, H3+
    double a19_11 = 0.0 - c * m.J(81);
    double a19_17 = 0.0;
    double a19_19 = 1.0 - c * m.J(82);
    double a19_20 = 0.0;
    double a19_22 = 0.0;
    double a19_24 = 0.0 - c * m.J(83);
    double a19_25 = 0.0 - c * m.J(84);
    double a19_26 = 0.0;
    double a19_27 = 0.0;
    double a19_29 = 0.0 - c * m.J(85);
    double a20_0 = 0.0 - c * m.J(86);
    double a20_20 = 1.0 - c * m.J(87);
    double a20_24 = 0.0 - c * m.J(88);
    double a20_26 = 0.0 - c * m.J(89);
    double a20_27 = 0.0 - c * m.J(90);
    double a20_28 = 0.0 - c * m.J(91);
    double a20_29 = 0.0 - c * m.J(92);
    a19_11 *= m.LU(35);
    a19_17 -= a19_11 * m.LU(36);
    a19_24 -= a19_11 * m.LU(37);
    a19_26 -= a19_11 * m.LU(38);
    a19_17 *= m.LU(65);
    a19_20 -= a19_17 * m.LU(66);
    a19_22 -= a19_17 * m.LU(67);
    a19_24 -= a19_17 * m.LU(68);
    a19_26 -= a19_17 * m.LU(69);
    a19_27 -= a19_17 * m.LU(70);
    a20_0 *= m.LU(0);
    a20_26 -= a20_0 * m.LU(1);
    m.LU(81) = a19_11;
    m.LU(82) = a19_17;
    m.LU(83) = 1.0 / a19_19;
    m.LU(84) = a19_20;
    m.LU(85) = a19_22;
    m.LU(86) = a19_24;
    m.LU(87) = a19_25;
    m.LU(88) = a19_26;
    m.LU(89) = a19_27;
    m.LU(90) = a19_29;
    m.LU(91) = a20_0;
    m.LU(92) = 1.0 / a20_20;
    m.LU(93) = a20_24;
    m.LU(94) = a20_26;
    m.LU(95) = a20_27;
    m.LU(96) = a20_28;
    m.LU(97) = a20_29;
  }
  P3D_FENCE();
  {  // level 3: H2O
    double a21_12 = 0.0 - c * m.J(93);
    double a21_14 = 0.0;
    double a21_16 = 0.0;
    double a21_20 = 0.0;
    double a21_21 = 1.0 - c * m.J(94);
    double a21_22 = 0.0 - c * m.J(95);
    double a21_23 = 0.0 - c * m.J(96);
    double a21_24 = 0.0 - c * m.J(97);
    double a21_26 = 0.0 - c * m.J(98);
    double a21_27 = 0.0 - c * m.J(99);
    double a21_28 = 0.0;
    double a21_29 = 0.0;
    a21_12 *= m.LU(39);
    a21_14 -= a21_12 * m.LU(40);
    a21_24 -= a21_12 * m.LU(41);
    a21_26 -= a21_12 * m.LU(42);
    a21_14 *= m.LU(51);
    a21_16 -= a21_14 * m.LU(52);
    a21_24 -= a21_14 * m.LU(53);
    a21_26 -= a21_14 * m.LU(54);
    a21_16 *= m.LU(60);
    a21_20 -= a21_16 * m.LU(61);
    a21_24 -= a21_16 * m.LU(62);
    a21_26 -= a21_16 * m.LU(63);
    a21_29 -= a21_16 * m.LU(64);
    a21_20 *= m.LU(92);
    a21_24 -= a21_20 * m.LU(93);
    a21_26 -= a21_20 * m.LU(94);
    a21_27 -= a21_20 * m.LU(95);
    a21_28 -= a21_20 * m.LU(96);
    a21_29 -= a21_20 * m.LU(97);
    m.LU(98) = a21_12;
    m.LU(99) = a21_14;
    m.LU(100) = a21_16;
    m.LU(101) = a21_20;
    m.LU(102) = 1.0 / a21_21;
    m.LU(103) = a21_22;
    m.LU(104) = a21_23;
    m.LU(105) = a21_24;
    m.LU(106) = a21_26;
    m.LU(107) = a21_27;
    m.LU(108) = a21_28;
    m.LU(109) = a21_29;
  }
  P3D_FENCE();
  {  // level 4: C+
    double a22_9 = 0.0 - c * m.J(100);
    double a22_15 = 0.0 - c * m.J(101);
    double a22_21 = 0.0 - c * m.J(102);
    double a22_22 = 1.0 - c * m.J(103);
    double a22_23 = 0.0 - c * m.J(104);
    double a22_24 = 0.0 - c * m.J(105);
    double a22_26 = 0.0 - c * m.J(106);
    double a22_27 = 0.0 - c * m.J(107);
    double a22_28 = 0.0 - c * m.J(108);
    double a22_29 = 0.0;
    a22_9 *= m.LU(28);
    a22_24 -= a22_9 * m.LU(29);
    a22_28 -= a22_9 * m.LU(30);
    a22_15 *= m.LU(55);
    a22_22 -= a22_15 * m.LU(56);
    a22_23 -= a22_15 * m.LU(57);
    a22_27 -= a22_15 * m.LU(58);
    a22_29 -= a22_15 * m.LU(59);
    a22_21 *= m.LU(102);
    a22_22 -= a22_21 * m.LU(103);
    a22_23 -= a22_21 * m.LU(104);
    a22_24 -= a22_21 * m.LU(105);
    a22_26 -= a22_21 * m.LU(106);
    a22_27 -= a22_21 * m.LU(107);
    a22_28 -= a22_21 * m.LU(108);
    a22_29 -= a22_21 * m.LU(109);
    m.LU(110) = a22_9;
    m.LU(111) = a22_15;
    m.LU(112) = a22_21;
    m.LU(113) = 1.0 / a22_22;
    m.LU(114) = a22_23;
    m.LU(115) = a22_24;
    m.LU(116) = a22_26;
    m.LU(117) = a22_27;
    m.LU(118) = a22_28;
    m.LU(119) = a22_29;
  }
  P3D_FENCE();
  {  // level 5: OH
    double a23_12 = 0.0 - c * m.J(109);
    double a23_14 = 0.0 - c * m.J(110);
    double a23_16 = 0.0;
    double a23_20 = 0.0;
    double a23_21 = 0.0 - c * m.J(111);
    double a23_22 = 0.0 - c * m.J(112);
    double a23_23 = 1.0 - c * m.J(113);
    double a23_24 = 0.0 - c * m.J(114);
    double a23_25 = 0.0 - c * m.J(115);
    double a23_26 = 0.0 - c * m.J(116);
    double a23_27 = 0.0 - c * m.J(117);
    double a23_28 = 0.0;
    double a23_29 = 0.0 - c * m.J(118);
    a23_12 *= m.LU(39);
    a23_14 -= a23_12 * m.LU(40);
    a23_24 -= a23_12 * m.LU(41);
    a23_26 -= a23_12 * m.LU(42);
    a23_14 *= m.LU(51);
    a23_16 -= a23_14 * m.LU(52);
    a23_24 -= a23_14 * m.LU(53);
    a23_26 -= a23_14 * m.LU(54);
    a23_16 *= m.LU(60);
    a23_20 -= a23_16 * m.LU(61);
    a23_24 -= a23_16 * m.LU(62);
    a23_26 -= a23_16 * m.LU(63);
    a23_29 -= a23_16 * m.LU(64);
    a23_20 *= m.LU(92);
    a23_24 -= a23_20 * m.LU(93);
    a23_26 -= a23_20 * m.LU(94);
    a23_27 -= a23_20 * m.LU(95);
    a23_28 -= a23_20 * m.LU(96);
    a23_29 -= a23_20 * m.LU(97);
    a23_21 *= m.LU(102);
    a23_22 -= a23_21 * m.LU(103);
    a23_23 -= a23_21 * m.LU(104);
    a23_24 -= a23_21 * m.LU(105);
    a23_26 -= a23_21 * m.LU(106);
    a23_27 -= a23_21 * m.LU(107);
    a23_28 -= a23_21 * m.LU(108);
    a23_29 -= a23_21 * m.LU(109);
    a23_22 *= m.LU(113);
    a23_23 -= a23_22 * m.LU(114);
    a23_24 -= a23_22 * m.LU(115);
    a23_26 -= a23_22 * m.LU(116);
    a23_27 -= a23_22 * m.LU(117);
    a23_28 -= a23_22 * m.LU(118);
    a23_29 -= a23_22 * m.LU(119);
    m.LU(120) = a23_12;
    m.LU(121) = a23_14;
    m.LU(122) = a23_16;
    m.LU(123) = a23_20;
    m.LU(124) = a23_21;
    m.LU(125) = a23_22;
    m.LU(126) = 1.0 / a23_23;
    m.LU(127) = a23_24;
    m.LU(128) = a23_25;
    m.LU(129) = a23_26;
    m.LU(130) = a23_27;
    m.LU(131) = a23_28;
    m.LU(132) = a23_29;
  }
  P3D_FENCE();
  {  // level 6: e-
    double a24_1 = 0.0 - c * m.J(119);
    double a24_2 = 0.0 - c * m.J(120);
    double a24_3 = 0.0 - c * m.J(121);
    double a24_4 = 0.0 - c * m.J(122);
    double a24_5 = 0.0 - c * m.J(123);
    double a24_6 = 0.0 - c * m.J(124);
    double a24_7 = 0.0 - c * m.J(125);
    double a24_8 = 0.0 - c * m.J(126);
    double a24_9 = 0.0 - c * m.J(127);
    double a24_11 = 0.0 - c * m.J(128);
    double a24_12 = 0.0 - c * m.J(129);
    double a24_13 = 0.0 - c * m.J(130);
    double a24_14 = 0.0 - c * m.J(131);
    double a24_16 = 0.0 - c * m.J(132);
    double a24_17 = 0.0 - c * m.J(133);
    double a24_18 = 0.0 - c * m.J(134);
    double a24_19 = 0.0 - c * m.J(135);
    double a24_20 = 0.0 - c * m.J(136);
    double a24_21 = 0.0;
    double a24_22 = 0.0 - c * m.J(137);
    double a24_23 = 0.0;
    double a24_24 = 1.0 - c * m.J(138);
    double a24_25 = 0.0 - c * m.J(139);
    double a24_26 = 0.0 - c * m.J(140);
    double a24_27 = 0.0 - c * m.J(141);
    double a24_28 = 0.0;
    double a24_29 = 0.0 - c * m.J(142);
    double a24_30 = 0.0 - c * m.J(143);
    a24_1 *= m.LU(2);
    a24_2 -= a24_1 * m.LU(3);
    a24_24 -= a24_1 * m.LU(4);
    a24_2 *= m.LU(6);
    a24_24 -= a24_2 * m.LU(7);
    a24_3 *= m.LU(8);
    a24_4 -= a24_3 * m.LU(9);
    a24_24 -= a24_3 * m.LU(10);
    a24_4 *= m.LU(12);
    a24_24 -= a24_4 * m.LU(13);
    a24_5 *= m.LU(14);
    a24_6 -= a24_5 * m.LU(15);
    a24_24 -= a24_5 * m.LU(16);
    a24_6 *= m.LU(18);
    a24_24 -= a24_6 * m.LU(19);
    a24_7 *= m.LU(20);
    a24_24 -= a24_7 * m.LU(21);
    a24_25 -= a24_7 * m.LU(22);
    a24_8 *= m.LU(23);
    a24_9 -= a24_8 * m.LU(24);
    a24_24 -= a24_8 * m.LU(25);
    a24_28 -= a24_8 * m.LU(26);
    a24_9 *= m.LU(28);
    a24_24 -= a24_9 * m.LU(29);
    a24_28 -= a24_9 * m.LU(30);
    a24_11 *= m.LU(35);
    a24_17 -= a24_11 * m.LU(36);
    a24_24 -= a24_11 * m.LU(37);
    a24_26 -= a24_11 * m.LU(38);
    a24_12 *= m.LU(39);
    a24_14 -= a24_12 * m.LU(40);
    a24_24 -= a24_12 * m.LU(41);
    a24_26 -= a24_12 * m.LU(42);
    a24_13 *= m.LU(44);
    a24_22 -= a24_13 * m.LU(45);
    a24_23 -= a24_13 * m.LU(46);
    a24_24 -= a24_13 * m.LU(47);
    a24_25 -= a24_13 * m.LU(48);
    a24_29 -= a24_13 * m.LU(49);
    a24_30 -= a24_13 * m.LU(50);
    a24_14 *= m.LU(51);
    a24_16 -= a24_14 * m.LU(52);
    a24_24 -= a24_14 * m.LU(53);
    a24_26 -= a24_14 * m.LU(54);
    a24_16 *= m.LU(60);
    a24_20 -= a24_16 * m.LU(61);
    a24_24 -= a24_16 * m.LU(62);
    a24_26 -= a24_16 * m.LU(63);
    a24_29 -= a24_16 * m.LU(64);
    a24_17 *= m.LU(65);
    a24_20 -= a24_17 * m.LU(66);
    a24_22 -= a24_17 * m.LU(67);
    a24_24 -= a24_17 * m.LU(68);
    a24_26 -= a24_17 * m.LU(69);
    a24_27 -= a24_17 * m.LU(70);
    a24_18 *= m.LU(71);
    a24_19 -= a24_18 * m.LU(72);
    a24_20 -= a24_18 * m.LU(73);
    a24_21 -= a24_18 * m.LU(74);
    a24_22 -= a24_18 * m.LU(75);
    a24_24 -= a24_18 * m.LU(76);
    a24_26 -= a24_18 * m.LU(77);
    a24_28 -= a24_18 * m.LU(78);
    a24_29 -= a24_18 * m.LU(79);
    a24_30 -= a24_18 * m.LU(80);
    a24_19 *= m.LU(83);
    a24_20 -= a24_19 * m.LU(84);
    a24_22 -= a24_19 * m.LU(85);
    a24_24 -= a24_19 * m.LU(86);
    a24_25 -= a24_19 * m.LU(87);
    a24_26 -= a24_19 * m.LU(88);
    a24_27 -= a24_19 * m.LU(89);
    a24_29 -= a24_19 * m.LU(90);
    a24_20 *= m.LU(92);
    a24_24 -= a24_20 * m.LU(93);
    a24_26 -= a24_20 * m.LU(94);
    a24_27 -= a24_20 * m.LU(95);
    a24_28 -= a24_20 * m.LU(96);
    a24_29 -= a24_20 * m.LU(97);
    a24_21 *= m.LU(102);
    a24_22 -= a24_21 * m.LU(103);
    a24_23 -= a24_21 * m.LU(104);
    a24_24 -= a24_21 * m.LU(105);
    a24_26 -= a24_21 * m.LU(106);
    a24_27 -= a24_21 * m.LU(107);
    a24_28 -= a24_21 * m.LU(108);
    a24_29 -= a24_21 * m.LU(109);
    a24_22 *= m.LU(113);
    a24_23 -= a24_22 * m.LU(114);
    a24_24 -= a24_22 * m.LU(115);
    a24_26 -= a24_22 * m.LU(116);
    a24_27 -= a24_22 * m.LU(117);
    a24_28 -= a24_22 * m.LU(118);
    a24_29 -= a24_22 * m.LU(119);
    a24_23 *= m.LU(126);
    a24_24 -= a24_23 * m.LU(127);
    a24_25 -= a24_23 * m.LU(128);
    a24_26 -= a24_23 * m.LU(129);
    a24_27 -= a24_23 * m.LU(130);
    a24_28 -= a24_23 * m.LU(131);
    a24_29 -= a24_23 * m.LU(132);
    m.LU(133) = a24_1;
    m.LU(134) = a24_2;
    m.LU(135) = a24_3;
    m.LU(136) = a24_4;
    m.LU(137) = a24_5;
    m.LU(138) = a24_6;
    m.LU(139) = a24_7;
    m.LU(140) = a24_8;
    m.LU(141) = a24_9;
    m.LU(142) = a24_11;
    m.LU(143) = a24_12;
    m.LU(144) = a24_13;
    m.LU(145) = a24_14;
    m.LU(146) = a24_16;
    m.LU(147) = a24_17;
    m.LU(148) = a24_18;
    m.LU(149) = a24_19;
    m.LU(150) = a24_20;
    m.LU(151) = a24_21;
    m.LU(152) = a24_22;
    m.LU(153) = a24_23;
    m.LU(154) = 1.0 / a24_24;
    m.LU(155) = a24_25;
    m.LU(156) = a24_26;
    m.LU(157) = a24_27;
    m.LU(158) = a24_28;
    m.LU(159) = a24_29;
    m.LU(160) = a24_30;
  }
  P3D_FENCE();
  {  // level 7: H
    double a25_0 = 0.0 - c * m.J(144);
    double a25_7 = 0.0 - c * m.J(145);
    double a25_10 = 0.0 - c * m.J(146);
    double a25_11 = 0.0 - c * m.J(147);
    double a25_12 = 0.0 - c * m.J(148);
    double a25_13 = 0.0 - c * m.J(149);
    double a25_14 = 0.0 - c * m.J(150);
    double a25_16 = 0.0 - c * m.J(151);
    double a25_17 = 0.0 - c * m.J(152);
    double a25_18 = 0.0 - c * m.J(153);
    double a25_19 = 0.0 - c * m.J(154);
    double a25_20 = 0.0 - c * m.J(155);
    double a25_21 = 0.0 - c * m.J(156);
    double a25_22 = 0.0 - c * m.J(157);
    double a25_23 = 0.0 - c * m.J(158);
    double a25_24 = 0.0 - c * m.J(159);
    double a25_25 = 1.0 - c * m.J(160);
    double a25_26 = 0.0 - c * m.J(161);
    double a25_27 = 0.0 - c * m.J(162);
    double a25_28 = 0.0;
    double a25_29 = 0.0 - c * m.J(163);
    double a25_30 = 0.0 - c * m.J(164);
    a25_0 *= m.LU(0);
    a25_26 -= a25_0 * m.LU(1);
    a25_7 *= m.LU(20);
    a25_24 -= a25_7 * m.LU(21);
    a25_25 -= a25_7 * m.LU(22);
    a25_10 *= m.LU(31);
    a25_13 -= a25_10 * m.LU(32);
    a25_25 -= a25_10 * m.LU(33);
    a25_29 -= a25_10 * m.LU(34);
    a25_11 *= m.LU(35);
    a25_17 -= a25_11 * m.LU(36);
    a25_24 -= a25_11 * m.LU(37);
    a25_26 -= a25_11 * m.LU(38);
    a25_12 *= m.LU(39);
    a25_14 -= a25_12 * m.LU(40);
    a25_24 -= a25_12 * m.LU(41);
    a25_26 -= a25_12 * m.LU(42);
    a25_13 *= m.LU(44);
    a25_22 -= a25_13 * m.LU(45);
    a25_23 -= a25_13 * m.LU(46);
    a25_24 -= a25_13 * m.LU(47);
    a25_25 -= a25_13 * m.LU(48);
    a25_29 -= a25_13 * m.LU(49);
    a25_30 -= a25_13 * m.LU(50);
    a25_14 *= m.LU(51);
    a25_16 -= a25_14 * m.LU(52);
    a25_24 -= a25_14 * m.LU(53);
    a25_26 -= a25_14 * m.LU(54);
    a25_16 *= m.LU(60);
    a25_20 -= a25_16 * m.LU(61);
    a25_24 -= a25_16 * m.LU(62);
    a25_26 -= a25_16 * m.LU(63);
    a25_29 -= a25_16 * m.LU(64);
    a25_17 *= m.LU(65);
    a25_20 -= a25_17 * m.LU(66);
    a25_22 -= a25_17 * m.LU(67);
    a25_24 -= a25_17 * m.LU(68);
    a25_26 -= a25_17 * m.LU(69);
    a25_27 -= a25_17 * m.LU(70);
    a25_18 *= m.LU(71);
    a25_19 -= a25_18 * m.LU(72);
    a25_20 -= a25_18 * m.LU(73);
    a25_21 -= a25_18 * m.LU(74);
    a25_22 -= a25_18 * m.LU(75);
    a25_24 -= a25_18 * m.LU(76);
    a25_26 -= a25_18 * m.LU(77);
    a25_28 -= a25_18 * m.LU(78);
    a25_29 -= a25_18 * m.LU(79);
    a25_30 -= a25_18 * m.LU(80);
    a25_19 *= m.LU(83);
    a25_20 -= a25_19 * m.LU(84);
    a25_22 -= a25_19 * m.LU(85);
    a25_24 -= a25_19 * m.LU(86);
    a25_25 -= a25_19 * m.LU(87);
    a25_26 -= a25_19 * m.LU(88);
    a25_27 -= a25_19 * m.LU(89);
    a25_29 -= a25_19 * m.LU(90);
    a25_20 *= m.LU(92);
    a25_24 -= a25_20 * m.LU(93);
    a25_26 -= a25_20 * m.LU(94);
    a25_27 -= a25_20 * m.LU(95);
    a25_28 -= a25_20 * m.LU(96);
    a25_29 -= a25_20 * m.LU(97);
    a25_21 *= m.LU(102);
    a25_22 -= a25_21 * m.LU(103);
    a25_23 -= a25_21 * m.LU(104);
    a25_24 -= a25_21 * m.LU(105);
    a25_26 -= a25_21 * m.LU(106);
    a25_27 -= a25_21 * m.LU(107);
    a25_28 -= a25_21 * m.LU(108);
    a25_29 -= a25_21 * m.LU(109);
    a25_22 *= m.LU(113);
    a25_23 -= a25_22 * m.LU(114);
    a25_24 -= a25_22 * m.LU(115);
    a25_26 -= a25_22 * m.LU(116);
    a25_27 -= a25_22 * m.LU(117);
    a25_28 -= a25_22 * m.LU(118);
    a25_29 -= a25_22 * m.LU(119);
    a25_23 *= m.LU(126);
    a25_24 -= a25_23 * m.LU(127);
    a25_25 -= a25_23 * m.LU(128);
    a25_26 -= a25_23 * m.LU(129);
    a25_27 -= a25_23 * m.LU(130);
    a25_28 -= a25_23 * m.LU(131);
    a25_29 -= a25_23 * m.LU(132);
    a25_24 *= m.LU(154);
    a25_25 -= a25_24 * m.LU(155);
    a25_26 -= a25_24 * m.LU(156);
    a25_27 -= a25_24 * m.LU(157);
    a25_28 -= a25_24 * m.LU(158);
    a25_29 -= a25_24 * m.LU(159);
    a25_30 -= a25_24 * m.LU(160);
    m.LU(161) = a25_0;
    m.LU(162) = a25_7;
    m.LU(163) = a25_10;
    m.LU(164) = a25_11;
    m.LU(165) = a25_12;
    m.LU(166) = a25_13;
    m.LU(167) = a25_14;
    m.LU(168) = a25_16;
    m.LU(169) = a25_17;
    m.LU(170) = a25_18;
    m.LU(171) = a25_19;
    m.LU(172) = a25_20;
    m.LU(173) = a25_21;
    m.LU(174) = a25_22;
    m.LU(175) = a25_23;
    m.LU(176) = a25_24;
    m.LU(177) = 1.0 / a25_25;
    m.LU(178) = a25_26;
    m.LU(179) = a25_27;
    m.LU(180) = a25_28;
    m.LU(181) = a25_29;
    m.LU(182) = a25_30;
  }
  P3D_FENCE();
  {  // level 8: H2
    double a26_0 = 0.0 - c * m.J(165);
    double a26_7 = 0.0 - c * m.J(166);
    double a26_14 = 0.0 - c * m.J(167);
    double a26_16 = 0.0 - c * m.J(168);
    double a26_17 = 0.0 - c * m.J(169);
    double a26_19 = 0.0 - c * m.J(170);
    double a26_20 = 0.0 - c * m.J(171);
    double a26_22 = 0.0 - c * m.J(172);
    double a26_23 = 0.0 - c * m.J(173);
    double a26_24 = 0.0 - c * m.J(174);
    double a26_25 = 0.0 - c * m.J(175);
    double a26_26 = 1.0 - c * m.J(176);
    double a26_27 = 0.0 - c * m.J(177);
    double a26_28 = 0.0 - c * m.J(178);
    double a26_29 = 0.0 - c * m.J(179);
    double a26_30 = 0.0 - c * m.J(180);
    a26_0 *= m.LU(0);
    a26_26 -= a26_0 * m.LU(1);
    a26_7 *= m.LU(20);
    a26_24 -= a26_7 * m.LU(21);
    a26_25 -= a26_7 * m.LU(22);
    a26_14 *= m.LU(51);
    a26_16 -= a26_14 * m.LU(52);
    a26_24 -= a26_14 * m.LU(53);
    a26_26 -= a26_14 * m.LU(54);
    a26_16 *= m.LU(60);
    a26_20 -= a26_16 * m.LU(61);
    a26_24 -= a26_16 * m.LU(62);
    a26_26 -= a26_16 * m.LU(63);
    a26_29 -= a26_16 * m.LU(64);
    a26_17 *= m.LU(65);
    a26_20 -= a26_17 * m.LU(66);
    a26_22 -= a26_17 * m.LU(67);
    a26_24 -= a26_17 * m.LU(68);
    a26_26 -= a26_17 * m.LU(69);
    a26_27 -= a26_17 * m.LU(70);
    a26_19 *= m.LU(83);
    a26_20 -= a26_19 * m.LU(84);
    a26_22 -= a26_19 * m.LU(85);
    a26_24 -= a26_19 * m.LU(86);
    a26_25 -= a26_19 * m.LU(87);
    a26_26 -= a26_19 * m.LU(88);
    a26_27 -= a26_19 * m.LU(89);
    a26_29 -= a26_19 * m.LU(90);
    a26_20 *= m.LU(92);
    a26_24 -= a26_20 * m.LU(93);
    a26_26 -= a26_20 * m.LU(94);
    a26_27 -= a26_20 * m.LU(95);
    a26_28 -= a26_20 * m.LU(96);
    a26_29 -= a26_20 * m.LU(97);
    a26_22 *= m.LU(113);
    a26_23 -= a26_22 * m.LU(114);
    a26_24 -= a26_22 * m.LU(115);
    a26_26 -= a26_22 * m.LU(116);
    a26_27 -= a26_22 * m.LU(117);
    a26_28 -= a26_22 * m.LU(118);
    a26_29 -= a26_22 * m.LU(119);
    a26_23 *= m.LU(126);
    a26_24 -= a26_23 * m.LU(127);
    a26_25 -= a26_23 * m.LU(128);
    a26_26 -= a26_23 * m.LU(129);
    a26_27 -= a26_23 * m.LU(130);
    a26_28 -= a26_23 * m.LU(131);
    a26_29 -= a26_23 * m.LU(132);
    a26_24 *= m.LU(154);
    a26_25 -= a26_24 * m.LU(155);
    a26_26 -= a26_24 * m.LU(156);
    a26_27 -= a26_24 * m.LU(157);
    a26_28 -= a26_24 * m.LU(158);
    a26_29 -= a26_24 * m.LU(159);
    a26_30 -= a26_24 * m.LU(160);
    a26_25 *= m.LU(177);
    a26_26 -= a26_25 * m.LU(178);
    a26_27 -= a26_25 * m.LU(179);
    a26_28 -= a26_25 * m.LU(180);
    a26_29 -= a26_25 * m.LU(181);
    a26_30 -= a26_25 * m.LU(182);
    m.LU(183) = a26_0;
    m.LU(184) = a26_7;
    m.LU(185) = a26_14;
    m.LU(186) = a26_16;
    m.LU(187) = a26_17;
    m.LU(188) = a26_19;
    m.LU(189) = a26_20;
    m.LU(190) = a26_22;
    m.LU(191) = a26_23;
    m.LU(192) = a26_24;
    m.LU(193) = a26_25;
    m.LU(194) = 1.0 / a26_26;
    m.LU(195) = a26_27;
    m.LU(196) = a26_28;
    m.LU(197) = a26_29;
    m.LU(198) = a26_30;
  }
  P3D_FENCE();
  {  // level 9: C
    double a27_11 = 0.0 - c * m.J(181);
    double a27_15 = 0.0 - c * m.J(182);
    double a27_17 = 0.0 - c * m.J(183);
    double a27_19 = 0.0 - c * m.J(184);
    double a27_20 = 0.0 - c * m.J(185);
    double a27_21 = 0.0 - c * m.J(186);
    double a27_22 = 0.0 - c * m.J(187);
    double a27_23 = 0.0 - c * m.J(188);
    double a27_24 = 0.0 - c * m.J(189);
    double a27_25 = 0.0 - c * m.J(190);
    double a27_26 = 0.0;
    double a27_27 = 1.0 - c * m.J(191);
    double a27_28 = 0.0 - c * m.J(192);
    double a27_29 = 0.0;
    double a27_30 = 0.0 - c * m.J(193);
    a27_11 *= m.LU(35);
    a27_17 -= a27_11 * m.LU(36);
    a27_24 -= a27_11 * m.LU(37);
    a27_26 -= a27_11 * m.LU(38);
    a27_15 *= m.LU(55);
    a27_22 -= a27_15 * m.LU(56);
    a27_23 -= a27_15 * m.LU(57);
    a27_27 -= a27_15 * m.LU(58);
    a27_29 -= a27_15 * m.LU(59);
    a27_17 *= m.LU(65);
    a27_20 -= a27_17 * m.LU(66);
    a27_22 -= a27_17 * m.LU(67);
    a27_24 -= a27_17 * m.LU(68);
    a27_26 -= a27_17 * m.LU(69);
    a27_27 -= a27_17 * m.LU(70);
    a27_19 *= m.LU(83);
    a27_20 -= a27_19 * m.LU(84);
    a27_22 -= a27_19 * m.LU(85);
    a27_24 -= a27_19 * m.LU(86);
    a27_25 -= a27_19 * m.LU(87);
    a27_26 -= a27_19 * m.LU(88);
    a27_27 -= a27_19 * m.LU(89);
    a27_29 -= a27_19 * m.LU(90);
    a27_20 *= m.LU(92);
    a27_24 -= a27_20 * m.LU(93);
    a27_26 -= a27_20 * m.LU(94);
    a27_27 -= a27_20 * m.LU(95);
    a27_28 -= a27_20 * m.LU(96);
    a27_29 -= a27_20 * m.LU(97);
    a27_21 *= m.LU(102);
    a27_22 -= a27_21 * m.LU(103);
    a27_23 -= a27_21 * m.LU(104);
    a27_24 -= a27_21 * m.LU(105);
    a27_26 -= a27_21 * m.LU(106);
    a27_27 -= a27_21 * m.LU(107);
    a27_28 -= a27_21 * m.LU(108);
    a27_29 -= a27_21 * m.LU(109);
    a27_22 *= m.LU(113);
    a27_23 -= a27_22 * m.LU(114);
    a27_24 -= a27_22 * m.LU(115);
    a27_26 -= a27_22 * m.LU(116);
    a27_27 -= a27_22 * m.LU(117);
    a27_28 -= a27_22 * m.LU(118);
    a27_29 -= a27_22 * m.LU(119);
    a27_23 *= m.LU(126);
    a27_24 -= a27_23 * m.LU(127);
    a27_25 -= a27_23 * m.LU(128);
    a27_26 -= a27_23 * m.LU(129);
    a27_27 -= a27_23 * m.LU(130);
    a27_28 -= a27_23 * m.LU(131);
    a27_29 -= a27_23 * m.LU(132);
    a27_24 *= m.LU(154);
    a27_25 -= a27_24 * m.LU(155);
    a27_26 -= a27_24 * m.LU(156);
    a27_27 -= a27_24 * m.LU(157);
    a27_28 -= a27_24 * m.LU(158);
    a27_29 -= a27_24 * m.LU(159);
    a27_30 -= a27_24 * m.LU(160);
    a27_25 *= m.LU(177);
    a27_26 -= a27_25 * m.LU(178);
    a27_27 -= a27_25 * m.LU(179);
    a27_28 -= a27_25 * m.LU(180);
    a27_29 -= a27_25 * m.LU(181);
    a27_30 -= a27_25 * m.LU(182);
    a27_26 *= m.LU(194);
    a27_27 -= a27_26 * m.LU(195);
    a27_28 -= a27_26 * m.LU(196);
    a27_29 -= a27_26 * m.LU(197);
    a27_30 -= a27_26 * m.LU(198);
    m.LU(199) = a27_11;
    m.LU(200) = a27_15;
    m.LU(201) = a27_17;
    m.LU(202) = a27_19;
    m.LU(203) = a27_20;
    m.LU(204) = a27_21;
    m.LU(205) = a27_22;
    m.LU(206) = a27_23;
    m.LU(207) = a27_24;
    m.LU(208) = a27_25;
    m.LU(209) = a27_26;
    m.LU(210) = 1.0 / a27_27;
    m.LU(211) = a27_28;
    m.LU(212) = a27_29;
    m.LU(213) = a27_30;
  }
  P3D_FENCE();
  {  // level 10: CO
    double a28_9 = 0.0 - c * m.J(194);
    double a28_15 = 0.0 - c * m.J(195);
    double a28_18 = 0.0 - c * m.J(196);
    double a28_19 = 0.0 - c * m.J(197);
    double a28_20 = 0.0 - c * m.J(198);
    double a28_21 = 0.0 - c * m.J(199);
    double a28_22 = 0.0 - c * m.J(200);
    double a28_23 = 0.0 - c * m.J(201);
    double a28_24 = 0.0 - c * m.J(202);
    double a28_25 = 0.0 - c * m.J(203);
    double a28_26 = 0.0;
    double a28_27 = 0.0 - c * m.J(204);
    double a28_28 = 1.0 - c * m.J(205);
    double a28_29 = 0.0 - c * m.J(206);
    double a28_30 = 0.0 - c * m.J(207);
    a28_9 *= m.LU(28);
    a28_24 -= a28_9 * m.LU(29);
    a28_28 -= a28_9 * m.LU(30);
    a28_15 *= m.LU(55);
    a28_22 -= a28_15 * m.LU(56);
    a28_23 -= a28_15 * m.LU(57);
    a28_27 -= a28_15 * m.LU(58);
    a28_29 -= a28_15 * m.LU(59);
    a28_18 *= m.LU(71);
    a28_19 -= a28_18 * m.LU(72);
    a28_20 -= a28_18 * m.LU(73);
    a28_21 -= a28_18 * m.LU(74);
    a28_22 -= a28_18 * m.LU(75);
    a28_24 -= a28_18 * m.LU(76);
    a28_26 -= a28_18 * m.LU(77);
    a28_28 -= a28_18 * m.LU(78);
    a28_29 -= a28_18 * m.LU(79);
    a28_30 -= a28_18 * m.LU(80);
    a28_19 *= m.LU(83);
    a28_20 -= a28_19 * m.LU(84);
    a28_22 -= a28_19 * m.LU(85);
    a28_24 -= a28_19 * m.LU(86);
    a28_25 -= a28_19 * m.LU(87);
    a28_26 -= a28_19 * m.LU(88);
    a28_27 -= a28_19 * m.LU(89);
    a28_29 -= a28_19 * m.LU(90);
    a28_20 *= m.LU(92);
    a28_24 -= a28_20 * m.LU(93);
    a28_26 -= a28_20 * m.LU(94);
    a28_27 -= a28_20 * m.LU(95);
    a28_28 -= a28_20 * m.LU(96);
    a28_29 -= a28_20 * m.LU(97);
    a28_21 *= m.LU(102);
    a28_22 -= a28_21 * m.LU(103);
    a28_23 -= a28_21 * m.LU(104);
    a28_24 -= a28_21 * m.LU(105);
    a28_26 -= a28_21 * m.LU(106);
    a28_27 -= a28_21 * m.LU(107);
    a28_28 -= a28_21 * m.LU(108);
    a28_29 -= a28_21 * m.LU(109);
    a28_22 *= m.LU(113);
    a28_23 -= a28_22 * m.LU(114);
    a28_24 -= a28_22 * m.LU(115);
    a28_26 -= a28_22 * m.LU(116);
    a28_27 -= a28_22 * m.LU(117);
    a28_28 -= a28_22 * m.LU(118);
    a28_29 -= a28_22 * m.LU(119);
    a28_23 *= m.LU(126);
    a28_24 -= a28_23 * m.LU(127);
    a28_25 -= a28_23 * m.LU(128);
    a28_26 -= a28_23 * m.LU(129);
    a28_27 -= a28_23 * m.LU(130);
    a28_28 -= a28_23 * m.LU(131);
    a28_29 -= a28_23 * m.LU(132);
    a28_24 *= m.LU(154);
    a28_25 -= a28_24 * m.LU(155);
    a28_26 -= a28_24 * m.LU(156);
    a28_27 -= a28_24 * m.LU(157);
    a28_28 -= a28_24 * m.LU(158);
    a28_29 -= a28_24 * m.LU(159);
    a28_30 -= a28_24 * m.LU(160);
    a28_25 *= m.LU(177);
    a28_26 -= a28_25 * m.LU(178);
    a28_27 -= a28_25 * m.LU(179);
    a28_28 -= a28_25 * m.LU(180);
    a28_29 -= a28_25 * m.LU(181);
    a28_30 -= a28_25 * m.LU(182);
    a28_26 *= m.LU(194);
    a28_27 -= a28_26 * m.LU(195);
    a28_28 -= a28_26 * m.LU(196);
    a28_29 -= a28_26 * m.LU(197);
    a28_30 -= a28_26 * m.LU(198);
    a28_27 *= m.LU(210);
    a28_28 -= a28_27 * m.LU(211);
    a28_29 -= a28_27 * m.LU(212);
    a28_30 -= a28_27 * m.LU(213);
    m.LU(214) = a28_9;
    m.LU(215) = a28_15;
    m.LU(216) = a28_18;
    m.LU(217) = a28_19;
    m.LU(218) = a28_20;
    m.LU(219) = a28_21;
    m.LU(220) = a28_22;
    m.LU(221) = a28_23;
    m.LU(222) = a28_24;
    m.LU(223) = a28_25;
    m.LU(224) = a28_26;
    m.LU(225) = a28_27;
    m.LU(226) = 1.0 / a28_28;
    m.LU(227) = a28_29;
    m.LU(228) = a28_30;
  }
  P3D_FENCE();
  {  // level 11: O
    double a29_9 = 0.0 - c * m.J(208);
    double a29_10 = 0.0 - c * m.J(209);
    double a29_13 = 0.0 - c * m.J(210);
    double a29_14 = 0.0 - c * m.J(211);
    double a29_15 = 0.0 - c * m.J(212);
    double a29_16 = 0.0 - c * m.J(213);
    double a29_19 = 0.0 - c * m.J(214);
    double a29_20 = 0.0 - c * m.J(215);
    double a29_22 = 0.0 - c * m.J(216);
    double a29_23 = 0.0 - c * m.J(217);
    double a29_24 = 0.0 - c * m.J(218);
    double a29_25 = 0.0 - c * m.J(219);
    double a29_26 = 0.0 - c * m.J(220);
    double a29_27 = 0.0 - c * m.J(221);
    double a29_28 = 0.0 - c * m.J(222);
    double a29_29 = 1.0 - c * m.J(223);
    double a29_30 = 0.0 - c * m.J(224);
    a29_9 *= m.LU(28);
    a29_24 -= a29_9 * m.LU(29);
    a29_28 -= a29_9 * m.LU(30);
    a29_10 *= m.LU(31);
    a29_13 -= a29_10 * m.LU(32);
    a29_25 -= a29_10 * m.LU(33);
    a29_29 -= a29_10 * m.LU(34);
    a29_13 *= m.LU(44);
    a29_22 -= a29_13 * m.LU(45);
    a29_23 -= a29_13 * m.LU(46);
    a29_24 -= a29_13 * m.LU(47);
    a29_25 -= a29_13 * m.LU(48);
    a29_29 -= a29_13 * m.LU(49);
    a29_30 -= a29_13 * m.LU(50);
    a29_14 *= m.LU(51);
    a29_16 -= a29_14 * m.LU(52);
    a29_24 -= a29_14 * m.LU(53);
    a29_26 -= a29_14 * m.LU(54);
    a29_15 *= m.LU(55);
    a29_22 -= a29_15 * m.LU(56);
    a29_23 -= a29_15 * m.LU(57);
    a29_27 -= a29_15 * m.LU(58);
    a29_29 -= a29_15 * m.LU(59);
    a29_16 *= m.LU(60);
    a29_20 -= a29_16 * m.LU(61);
    a29_24 -= a29_16 * m.LU(62);
    a29_26 -= a29_16 * m.LU(63);
    a29_29 -= a29_16 * m.LU(64);
    a29_19 *= m.LU(83);
    a29_20 -= a29_19 * m.LU(84);
    a29_22 -= a29_19 * m.LU(85);
    a29_24 -= a29_19 * m.LU(86);
    a29_25 -= a29_19 * m.LU(87);
    a29_26 -= a29_19 * m.LU(88);
    a29_27 -= a29_19 * m.LU(89);
    a29_29 -= a29_19 * m.LU(90);
    a29_20 *= m.LU(92);
    a29_24 -= a29_20 * m.LU(93);
    a29_26 -= a29_20 * m.LU(94);
    a29_27 -= a29_20 * m.LU(95);
    a29_28 -= a29_20 * m.LU(96);
    a29_29 -= a29_20 * m.LU(97);
    a29_22 *= m.LU(113);
    a29_23 -= a29_22 * m.LU(114);
    a29_24 -= a29_22 * m.LU(115);
    a29_26 -= a29_22 * m.LU(116);
    a29_27 -= a29_22 * m.LU(117);
    a29_28 -= a29_22 * m.LU(118);
    a29_29 -= a29_22 * m.LU(119);
    a29_23 *= m.LU(126);
    a29_24 -= a29_23 * m.LU(127);
    a29_25 -= a29_23 * m.LU(128);
    a29_26 -= a29_23 * m.LU(129);
    a29_27 -= a29_23 * m.LU(130);
    a29_28 -= a29_23 * m.LU(131);
    a29_29 -= a29_23 * m.LU(132);
    a29_24 *= m.LU(154);
    a29_25 -= a29_24 * m.LU(155);
    a29_26 -= a29_24 * m.LU(156);
    a29_27 -= a29_24 * m.LU(157);
    a29_28 -= a29_24 * m.LU(158);
    a29_29 -= a29_24 * m.LU(159);
    a29_30 -= a29_24 * m.LU(160);
    a29_25 *= m.LU(177);
    a29_26 -= a29_25 * m.LU(178);
    a29_27 -= a29_25 * m.LU(179);
    a29_28 -= a29_25 * m.LU(180);
    a29_29 -= a29_25 * m.LU(181);
    a29_30 -= a29_25 * m.LU(182);
    a29_26 *= m.LU(194);
    a29_27 -= a29_26 * m.LU(195);
    a29_28 -= a29_26 * m.LU(196);
    a29_29 -= a29_26 * m.LU(197);
    a29_30 -= a29_26 * m.LU(198);
    a29_27 *= m.LU(210);
    a29_28 -= a29_27 * m.LU(211);
    a29_29 -= a29_27 * m.LU(212);
    a29_30 -= a29_27 * m.LU(213);
    a29_28 *= m.LU(226);
    a29_29 -= a29_28 * m.LU(227);
    a29_30 -= a29_28 * m.LU(228);
    m.LU(229) = a29_9;
    m.LU(230) = a29_10;
    m.LU(231) = a29_13;
    m.LU(232) = a29_14;
    m.LU(233) = a29_15;
    m.LU(234) = a29_16;
    m.LU(235) = a29_19;
    m.LU(236) = a29_20;
    m.LU(237) = a29_22;
    m.LU(238) = a29_23;
    m.LU(239) = a29_24;
    m.LU(240) = a29_25;
    m.LU(241) = a29_26;
    m.LU(242) = a29_27;
    m.LU(243) = a29_28;
    m.LU(244) = 1.0 / a29_29;
    m.LU(245) = a29_30;
  }
  P3D_FENCE();
  {  // level 12: CO+
    double a30_15 = 0.0 - c * m.J(225);
    double a30_18 = 0.0 - c * m.J(226);
    double a30_19 = 0.0;
    double a30_20 = 0.0;
    double a30_21 = 0.0;
    double a30_22 = 0.0 - c * m.J(227);
    double a30_23 = 0.0;
    double a30_24 = 0.0 - c * m.J(228);
    double a30_25 = 0.0 - c * m.J(229);
    double a30_26 = 0.0 - c * m.J(230);
    double a30_27 = 0.0;
    double a30_28 = 0.0;
    double a30_29 = 0.0;
    double a30_30 = 1.0 - c * m.J(231);
    a30_15 *= m.LU(55);
    a30_22 -= a30_15 * m.LU(56);
    a30_23 -= a30_15 * m.LU(57);
    a30_27 -= a30_15 * m.LU(58);
    a30_29 -= a30_15 * m.LU(59);
    a30_18 *= m.LU(71);
    a30_19 -= a30_18 * m.LU(72);
    a30_20 -= a30_18 * m.LU(73);
    a30_21 -= a30_18 * m.LU(74);
    a30_22 -= a30_18 * m.LU(75);
    a30_24 -= a30_18 * m.LU(76);
    a30_26 -= a30_18 * m.LU(77);
    a30_28 -= a30_18 * m.LU(78);
    a30_29 -= a30_18 * m.LU(79);
    a30_30 -= a30_18 * m.LU(80);
    a30_19 *= m.LU(83);
    a30_20 -= a30_19 * m.LU(84);
    a30_22 -= a30_19 * m.LU(85);
    a30_24 -= a30_19 * m.LU(86);
    a30_25 -= a30_19 * m.LU(87);
    a30_26 -= a30_19 * m.LU(88);
    a30_27 -= a30_19 * m.LU(89);
    a30_29 -= a30_19 * m.LU(90);
    a30_20 *= m.LU(92);
    a30_24 -= a30_20 * m.LU(93);
    a30_26 -= a30_20 * m.LU(94);
    a30_27 -= a30_20 * m.LU(95);
    a30_28 -= a30_20 * m.LU(96);
    a30_29 -= a30_20 * m.LU(97);
    a30_21 *= m.LU(102);
    a30_22 -= a30_21 * m.LU(103);
    a30_23 -= a30_21 * m.LU(104);
    a30_24 -= a30_21 * m.LU(105);
    a30_26 -= a30_21 * m.LU(106);
    a30_27 -= a30_21 * m.LU(107);
    a30_28 -= a30_21 * m.LU(108);
    a30_29 -= a30_21 * m.LU(109);
    a30_22 *= m.LU(113);
    a30_23 -= a30_22 * m.LU(114);
    a30_24 -= a30_22 * m.LU(115);
    a30_26 -= a30_22 * m.LU(116);
    a30_27 -= a30_22 * m.LU(117);
    a30_28 -= a30_22 * m.LU(118);
    a30_29 -= a30_22 * m.LU(119);
    a30_23 *= m.LU(126);
    a30_24 -= a30_23 * m.LU(127);
    a30_25 -= a30_23 * m.LU(128);
    a30_26 -= a30_23 * m.LU(129);
    a30_27 -= a30_23 * m.LU(130);
    a30_28 -= a30_23 * m.LU(131);
    a30_29 -= a30_23 * m.LU(132);
    a30_24 *= m.LU(154);
    a30_25 -= a30_24 * m.LU(155);
    a30_26 -= a30_24 * m.LU(156);
    a30_27 -= a30_24 * m.LU(157);
    a30_28 -= a30_24 * m.LU(158);
    a30_29 -= a30_24 * m.LU(159);
    a30_30 -= a30_24 * m.LU(160);
    a30_25 *= m.LU(177);
    a30_26 -= a30_25 * m.LU(178);
    a30_27 -= a30_25 * m.LU(179);
    a30_28 -= a30_25 * m.LU(180);
    a30_29 -= a30_25 * m.LU(181);
    a30_30 -= a30_25 * m.LU(182);
    a30_26 *= m.LU(194);
    a30_27 -= a30_26 * m.LU(195);
    a30_28 -= a30_26 * m.LU(196);
    a30_29 -= a30_26 * m.LU(197);
    a30_30 -= a30_26 * m.LU(198);
    a30_27 *= m.LU(210);
    a30_28 -= a30_27 * m.LU(211);
    a30_29 -= a30_27 * m.LU(212);
    a30_30 -= a30_27 * m.LU(213);
    a30_28 *= m.LU(226);
    a30_29 -= a30_28 * m.LU(227);
    a30_30 -= a30_28 * m.LU(228);
    a30_29 *= m.LU(244);
    a30_30 -= a30_29 * m.LU(245);
    m.LU(246) = a30_15;
    m.LU(247) = a30_18;
    m.LU(248) = a30_19;
    m.LU(249) = a30_20;
    m.LU(250) = a30_21;
    m.LU(251) = a30_22;
    m.LU(252) = a30_23;
    m.LU(253) = a30_24;
    m.LU(254) = a30_25;
    m.LU(255) = a30_26;
    m.LU(256) = a30_27;
    m.LU(257) = a30_28;
    m.LU(258) = a30_29;
    m.LU(259) = 1.0 / a30_30;
  }
  P3D_FENCE();
}

// Solve M x = b in place (b in species order).
template <class Slot>
P3D_HD P3D_INLINE void lu_solve(const Slot &m, double (&b)[kNS]) {
  b[24] -= m.LU(5) * b[23];
  b[26] -= m.LU(11) * b[25];
  b[28] -= m.LU(17) * b[27];
  P3D_FENCE();
  b[30] -= m.LU(27) * b[29];
  b[3] -= m.LU(43) * b[14];
  P3D_FENCE();
  b[10] -= m.LU(81) * b[12];
  b[10] -= m.LU(82) * b[11];
  b[6] -= m.LU(91) * b[5];
  b[17] -= m.LU(98) * b[19];
  b[17] -= m.LU(99) * b[18];
  b[17] -= m.LU(100) * b[16];
  b[17] -= m.LU(101) * b[6];
  b[7] -= m.LU(110) * b[30];
  b[7] -= m.LU(111) * b[20];
  b[7] -= m.LU(112) * b[17];
  b[15] -= m.LU(120) * b[19];
  b[15] -= m.LU(121) * b[18];
  b[15] -= m.LU(122) * b[16];
  b[15] -= m.LU(123) * b[6];
  b[15] -= m.LU(124) * b[17];
  b[15] -= m.LU(125) * b[7];
  P3D_FENCE();
  b[0] -= m.LU(133) * b[23];
  b[0] -= m.LU(134) * b[24];
  b[0] -= m.LU(135) * b[25];
  b[0] -= m.LU(136) * b[26];
  b[0] -= m.LU(137) * b[27];
  b[0] -= m.LU(138) * b[28];
  b[0] -= m.LU(139) * b[4];
  b[0] -= m.LU(140) * b[29];
  b[0] -= m.LU(141) * b[30];
  b[0] -= m.LU(142) * b[12];
  b[0] -= m.LU(143) * b[19];
  b[0] -= m.LU(144) * b[3];
  b[0] -= m.LU(145) * b[18];
  b[0] -= m.LU(146) * b[16];
  b[0] -= m.LU(147) * b[11];
  b[0] -= m.LU(148) * b[22];
  b[0] -= m.LU(149) * b[10];
  b[0] -= m.LU(150) * b[6];
  b[0] -= m.LU(151) * b[17];
  b[0] -= m.LU(152) * b[7];
  b[0] -= m.LU(153) * b[15];
  b[1] -= m.LU(161) * b[5];
  b[1] -= m.LU(162) * b[4];
  b[1] -= m.LU(163) * b[14];
  b[1] -= m.LU(164) * b[12];
  b[1] -= m.LU(165) * b[19];
  b[1] -= m.LU(166) * b[3];
  b[1] -= m.LU(167) * b[18];
  b[1] -= m.LU(168) * b[16];
  b[1] -= m.LU(169) * b[11];
  b[1] -= m.LU(170) * b[22];
  b[1] -= m.LU(171) * b[10];
  b[1] -= m.LU(172) * b[6];
  b[1] -= m.LU(173) * b[17];
  b[1] -= m.LU(174) * b[7];
  b[1] -= m.LU(175) * b[15];
  b[1] -= m.LU(176) * b[0];
  b[2] -= m.LU(183) * b[5];
  b[2] -= m.LU(184) * b[4];
  b[2] -= m.LU(185) * b[18];
  b[2] -= m.LU(186) * b[16];
  b[2] -= m.LU(187) * b[11];
  b[2] -= m.LU(188) * b[10];
  b[2] -= m.LU(189) * b[6];
  b[2] -= m.LU(190) * b[7];
  b[2] -= m.LU(191) * b[15];
  b[2] -= m.LU(192) * b[0];
  b[2] -= m.LU(193) * b[1];
  b[8] -= m.LU(199) * b[12];
  b[8] -= m.LU(200) * b[20];
  b[8] -= m.LU(201) * b[11];
  b[8] -= m.LU(202) * b[10];
  b[8] -= m.LU(203) * b[6];
  b[8] -= m.LU(204) * b[17];
  b[8] -= m.LU(205) * b[7];
  b[8] -= m.LU(206) * b[15];
  b[8] -= m.LU(207) * b[0];
  b[8] -= m.LU(208) * b[1];
  b[8] -= m.LU(209) * b[2];
  b[9] -= m.LU(214) * b[30];
  b[9] -= m.LU(215) * b[20];
  b[9] -= m.LU(216) * b[22];
  b[9] -= m.LU(217) * b[10];
  b[9] -= m.LU(218) * b[6];
  b[9] -= m.LU(219) * b[17];
  b[9] -= m.LU(220) * b[7];
  b[9] -= m.LU(221) * b[15];
  b[9] -= m.LU(222) * b[0];
  b[9] -= m.LU(223) * b[1];
  b[9] -= m.LU(224) * b[2];
  b[9] -= m.LU(225) * b[8];
  b[13] -= m.LU(229) * b[30];
  b[13] -= m.LU(230) * b[14];
  b[13] -= m.LU(231) * b[3];
  b[13] -= m.LU(232) * b[18];
  b[13] -= m.LU(233) * b[20];
  b[13] -= m.LU(234) * b[16];
  b[13] -= m.LU(235) * b[10];
  b[13] -= m.LU(236) * b[6];
  b[13] -= m.LU(237) * b[7];
  b[13] -= m.LU(238) * b[15];
  b[13] -= m.LU(239) * b[0];
  b[13] -= m.LU(240) * b[1];
  b[13] -= m.LU(241) * b[2];
  b[13] -= m.LU(242) * b[8];
  b[13] -= m.LU(243) * b[9];
  b[21] -= m.LU(246) * b[20];
  b[21] -= m.LU(247) * b[22];
  b[21] -= m.LU(248) * b[10];
  b[21] -= m.LU(249) * b[6];
  b[21] -= m.LU(250) * b[17];
  b[21] -= m.LU(251) * b[7];
  b[21] -= m.LU(252) * b[15];
  b[21] -= m.LU(253) * b[0];
  b[21] -= m.LU(254) * b[1];
  b[21] -= m.LU(255) * b[2];
  b[21] -= m.LU(256) * b[8];
  b[21] -= m.LU(257) * b[9];
  b[21] -= m.LU(258) * b[13];
  P3D_FENCE();
  b[21] *= m.LU(259);
  b[13] -= m.LU(245) * b[21];
  b[13] *= m.LU(244);
  b[9] -= m.LU(228) * b[21];
  b[9] -= m.LU(227) * b[13];
  b[9] *= m.LU(226);
  b[8] -= m.LU(213) * b[21];
  b[8] -= m.LU(212) * b[13];
  b[8] -= m.LU(211) * b[9];
  b[8] *= m.LU(210);
  b[2] -= m.LU(198) * b[21];
  b[2] -= m.LU(197) * b[13];
  b[2] -= m.LU(196) * b[9];
  b[2] -= m.LU(195) * b[8];
  b[2] *= m.LU(194);
  b[1] -= m.LU(182) * b[21];
  b[1] -= m.LU(181) * b[13];
  b[1] -= m.LU(180) * b[9];
  b[1] -= m.LU(179) * b[8];
  b[1] -= m.LU(178) * b[2];
  b[1] *= m.LU(177);
  b[0] -= m.LU(160) * b[21];
  b[0] -= m.LU(159) * b[13];
  b[0] -= m.LU(158) * b[9];
  b[0] -= m.LU(157) * b[8];
  b[0] -= m.LU(156) * b[2];
  b[0] -= m.LU(155) * b[1];
  b[0] *= m.LU(154);
  P3D_FENCE();
  b[15] -= m.LU(132) * b[13];
  b[15] -= m.LU(131) * b[9];
  b[15] -= m.LU(130) * b[8];
  b[15] -= m.LU(129) * b[2];
  b[15] -= m.LU(128) * b[1];
  b[15] -= m.LU(127) * b[0];
  b[15] *= m.LU(126);
  b[7] -= m.LU(119) * b[13];
  b[7] -= m.LU(118) * b[9];
  b[7] -= m.LU(117) * b[8];
  b[7] -= m.LU(116) * b[2];
  b[7] -= m.LU(115) * b[0];
  b[7] -= m.LU(114) * b[15];
  b[7] *= m.LU(113);
  b[17] -= m.LU(109) * b[13];
  b[17] -= m.LU(108) * b[9];
  b[17] -= m.LU(107) * b[8];
  b[17] -= m.LU(106) * b[2];
  b[17] -= m.LU(105) * b[0];
  b[17] -= m.LU(104) * b[15];
  b[17] -= m.LU(103) * b[7];
  b[17] *= m.LU(102);
  b[6] -= m.LU(97) * b[13];
  b[6] -= m.LU(96) * b[9];
  b[6] -= m.LU(95) * b[8];
  b[6] -= m.LU(94) * b[2];
  b[6] -= m.LU(93) * b[0];
  b[6] *= m.LU(92);
  b[10] -= m.LU(90) * b[13];
  b[10] -= m.LU(89) * b[8];
  b[10] -= m.LU(88) * b[2];
  b[10] -= m.LU(87) * b[1];
  b[10] -= m.LU(86) * b[0];
  b[10] -= m.LU(85) * b[7];
  b[10] -= m.LU(84) * b[6];
  b[10] *= m.LU(83);
  b[22] -= m.LU(80) * b[21];
  b[22] -= m.LU(79) * b[13];
  b[22] -= m.LU(78) * b[9];
  b[22] -= m.LU(77) * b[2];
  b[22] -= m.LU(76) * b[0];
  b[22] -= m.LU(75) * b[7];
  b[22] -= m.LU(74) * b[17];
  b[22] -= m.LU(73) * b[6];
  b[22] -= m.LU(72) * b[10];
  b[22] *= m.LU(71);
  b[11] -= m.LU(70) * b[8];
  b[11] -= m.LU(69) * b[2];
  b[11] -= m.LU(68) * b[0];
  b[11] -= m.LU(67) * b[7];
  b[11] -= m.LU(66) * b[6];
  b[11] *= m.LU(65);
  b[16] -= m.LU(64) * b[13];
  b[16] -= m.LU(63) * b[2];
  b[16] -= m.LU(62) * b[0];
  b[16] -= m.LU(61) * b[6];
  b[16] *= m.LU(60);
  P3D_FENCE();
  b[20] -= m.LU(59) * b[13];
  b[20] -= m.LU(58) * b[8];
  b[20] -= m.LU(57) * b[15];
  b[20] -= m.LU(56) * b[7];
  b[20] *= m.LU(55);
  b[18] -= m.LU(54) * b[2];
  b[18] -= m.LU(53) * b[0];
  b[18] -= m.LU(52) * b[16];
  b[18] *= m.LU(51);
  b[3] -= m.LU(50) * b[21];
  b[3] -= m.LU(49) * b[13];
  b[3] -= m.LU(48) * b[1];
  b[3] -= m.LU(47) * b[0];
  b[3] -= m.LU(46) * b[15];
  b[3] -= m.LU(45) * b[7];
  b[3] *= m.LU(44);
  b[19] -= m.LU(42) * b[2];
  b[19] -= m.LU(41) * b[0];
  b[19] -= m.LU(40) * b[18];
  b[19] *= m.LU(39);
  b[12] -= m.LU(38) * b[2];
  b[12] -= m.LU(37) * b[0];
  b[12] -= m.LU(36) * b[11];
  b[12] *= m.LU(35);
  b[14] -= m.LU(34) * b[13];
  b[14] -= m.LU(33) * b[1];
  b[14] -= m.LU(32) * b[3];
  b[14] *= m.LU(31);
  b[30] -= m.LU(30) * b[9];
  b[30] -= m.LU(29) * b[0];
  b[30] *= m.LU(28);
  b[29] -= m.LU(26) * b[9];
  b[29] -= m.LU(25) * b[0];
  b[29] -= m.LU(24) * b[30];
  b[29] *= m.LU(23);
  P3D_FENCE();
  b[4] -= m.LU(22) * b[1];
  b[4] -= m.LU(21) * b[0];
  b[4] *= m.LU(20);
  b[28] -= m.LU(19) * b[0];
  b[28] *= m.LU(18);
  b[27] -= m.LU(16) * b[0];
  b[27] -= m.LU(15) * b[28];
  b[27] *= m.LU(14);
  b[26] -= m.LU(13) * b[0];
  b[26] *= m.LU(12);
  b[25] -= m.LU(10) * b[0];
  b[25] -= m.LU(9) * b[26];
  b[25] *= m.LU(8);
  b[24] -= m.LU(7) * b[0];
  b[24] *= m.LU(6);
  b[23] -= m.LU(4) * b[0];
  b[23] -= m.LU(3) * b[24];
  b[23] *= m.LU(2);
  b[5] -= m.LU(1) * b[2];
  b[5] *= m.LU(0);
  P3D_FENCE();
}

}  // namespace lane
}  // namespace p3d
