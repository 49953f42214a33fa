// Generated by prism3d_amd/gen_lane.py from the network structure of
// prism3d_amd/network.py.  Do not edit; rerun the generator.
//
// Register LU of M = I - c J for 31 species, one cell per wavefront, lane =
// species: the no-pivot factorisation in the minimum-Markowitz order of
// lane_net.h, right-looking, broadcasting only structural U entries.
#pragma once

namespace p3d {
namespace wave {

constexpr int kNS = 31;
constexpr int kNR = 75;  // reactions
constexpr int kNNZ = 232;  // structural entries of the Jacobian
constexpr int kPerm[kNS] = {5, 23, 24, 25, 26, 27, 28, 4, 29, 30, 14, 12, 19, 3, 18, 20, 16, 11, 22, 10, 6, 17, 7, 15, 0, 1, 2, 8, 9, 13, 21};
// per species (lane): bit k set where its row has a structural multiplier
// (kLMask) or upper entry (kUMask) at elimination position k
#define P3D_WAVE_LMASK {0x00ff7bfeu, 0x01ff7c81u, 0x03db4081u, 0x00000400u, 0x00000000u, 0x00000000u, 0x00000001u, 0x00208200u, 0x07fa8800u, 0x0ffc8200u, 0x00020800u, 0x00000000u, 0x00000000u, 0x1fd9e600u, 0x00000000u, 0x00715000u, 0x00000000u, 0x00115000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x3ffc8000u, 0x00000000u, 0x00000000u, 0x00000002u, 0x00000000u, 0x00000008u, 0x00000000u, 0x00000020u, 0x00000000u, 0x00000100u}
#define P3D_WAVE_UMASK {0x7e000000u, 0x7c000000u, 0x78000000u, 0x63c00000u, 0x03000000u, 0x04000000u, 0x3d000000u, 0x3d800000u, 0x70000000u, 0x60000000u, 0x2f500000u, 0x0d500000u, 0x05020000u, 0x40000000u, 0x22002000u, 0x3f000000u, 0x25100000u, 0x3dc00000u, 0x05010000u, 0x05004000u, 0x28c00000u, 0x00000000u, 0x75780000u, 0x01000004u, 0x01000000u, 0x01000010u, 0x01000000u, 0x01000040u, 0x01000000u, 0x11000200u, 0x11000000u}

// RHS pieces: lane l sums piece l, at most kPieceLen entries of one species'
// row (byte q of the two words: reaction << 1 | product, 0xff unused); the
// pieces of species s are lanes kRowPiece[s] .. kRowPiece[s + 1] - 1
constexpr int kPieceLen = 6;
#define P3D_WAVE_PIECE_W0 {0x0a080705u, 0x2e2c2010u, 0x4d4a4944u, 0x6f5c5952u, 0x8e8c8483u, 0xffffff94u, 0x03030000u, 0x19170c0au, 0x2423211fu, 0x312f2d2du, 0x3f3d3936u, 0x54454545u, 0x7a716d6bu, 0x87878181u, 0x8b8b8989u, 0xffff9593u, 0x0d040201u, 0x342a2827u, 0x4743403bu, 0x8b8a8a88u, 0x22150807u, 0xffffff62u, 0xffff0c0bu, 0xffff4005u, 0x42413a26u, 0x1614100fu, 0xff786f5eu, 0x1813110eu, 0x726e695fu, 0xff868581u, 0x1b191512u, 0x73685a55u, 0x7f7c7a1au, 0x47201e1du, 0xff807e1fu, 0x25221a13u, 0x695b3734u, 0x77757573u, 0x93918582u, 0xffff2423u, 0x302d1814u, 0x6c6b3836u, 0xff902827u, 0x39322f16u, 0x94922a29u, 0xff2e2c2bu, 0x76747271u, 0x7956543fu, 0x3e3c3b17u, 0xff614b48u, 0xff604a49u, 0x8f654f4cu, 0x8e644e4du, 0x8d675350u, 0x8c665251u, 0xff5d5b58u, 0xff5c5a59u, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu}
#define P3D_WAVE_PIECE_W1 {0xffff0f0du, 0xffff423cu, 0xffff514eu, 0xffff807eu, 0xffff9290u, 0xffffffffu, 0xffff0906u, 0xffff1d1bu, 0xffff2b29u, 0xffff3533u, 0xffff4341u, 0xffff6357u, 0xffff7f7du, 0xffff8988u, 0xffff9391u, 0xffffffffu, 0xffff1e1cu, 0xffff3837u, 0xffff7b56u, 0xffffffffu, 0xffff5525u, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffff4644u, 0xffff5b1cu, 0xffffffffu, 0xffff4621u, 0xffff7d7bu, 0xffffffffu, 0xffff3d3au, 0xffffff87u, 0xffffff82u, 0xffffffffu, 0xffffffffu, 0xffff3126u, 0xffff706du, 0xffff7977u, 0xffffffffu, 0xffffffffu, 0xffff3533u, 0xffff9570u, 0xffffffffu, 0xffff866au, 0xffffffffu, 0xffffffffu, 0xffffff78u, 0xffffff84u, 0xffff8357u, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu}
#define P3D_WAVE_ROW_PIECE {0, 6, 16, 20, 22, 23, 24, 25, 27, 30, 32, 33, 34, 35, 39, 40, 42, 43, 44, 45, 46, 47, 48, 49, 50, 51, 52, 53, 54, 55, 56, 57}
constexpr int kRowPieceMax = 10;  // pieces of the longest row

// a: the lane's row of M (positions); lm: the lane's kLMask entry; pos, lev:
// its elimination position and P3D_WAVE_LEVELS entry.  Afterwards a[k] holds
// the multiplier l_ik below the diagonal, 1 / u_ii on it and u_ij above it
// (structural entries).  At each dependency level the level's pivots invert
// their diagonals in place at once, then the pivot rows are broadcast
// together and applied in position order.  A: an accessor of the lane's row
// in LDS (a[pos] indexes it at run time; bc reads the pivot row there).
template <class A, class BC>
__device__ __forceinline__ void lu_factor(A &a, unsigned lm, int lane, const BC &bc,
                                          int pos, int lev) {
  const int lvk = lev >> 16;
  {  // level 1: H2+, S, Si, Fe, H-, He, O+, CH2+, H3O+, O2, HCO+
    if (lvk == 1) a[pos] = 1.0 / a[pos];
    const double inv0 = bc(a[0], 5);
    const double u0_26 = bc(a[26], 5);
    const double inv1 = bc(a[1], 23);
    const double u1_2 = bc(a[2], 23);
    const double u1_24 = bc(a[24], 23);
    const double inv3 = bc(a[3], 25);
    const double u3_4 = bc(a[4], 25);
    const double u3_24 = bc(a[24], 25);
    const double inv5 = bc(a[5], 27);
    const double u5_6 = bc(a[6], 27);
    const double u5_24 = bc(a[24], 27);
    const double inv7 = bc(a[7], 4);
    const double u7_24 = bc(a[24], 4);
    const double u7_25 = bc(a[25], 4);
    const double inv8 = bc(a[8], 29);
    const double u8_9 = bc(a[9], 29);
    const double u8_24 = bc(a[24], 29);
    const double u8_28 = bc(a[28], 29);
    const double inv10 = bc(a[10], 14);
    const double u10_13 = bc(a[13], 14);
    const double u10_25 = bc(a[25], 14);
    const double u10_29 = bc(a[29], 14);
    const double inv11 = bc(a[11], 12);
    const double u11_17 = bc(a[17], 12);
    const double u11_24 = bc(a[24], 12);
    const double u11_26 = bc(a[26], 12);
    const double inv12 = bc(a[12], 19);
    const double u12_14 = bc(a[14], 19);
    const double u12_24 = bc(a[24], 19);
    const double u12_26 = bc(a[26], 19);
    const double inv15 = bc(a[15], 20);
    const double u15_22 = bc(a[22], 20);
    const double u15_23 = bc(a[23], 20);
    const double u15_27 = bc(a[27], 20);
    const double u15_29 = bc(a[29], 20);
    const double inv18 = bc(a[18], 22);
    const double u18_19 = bc(a[19], 22);
    const double u18_20 = bc(a[20], 22);
    const double u18_21 = bc(a[21], 22);
    const double u18_22 = bc(a[22], 22);
    const double u18_24 = bc(a[24], 22);
    const double u18_26 = bc(a[26], 22);
    const double u18_28 = bc(a[28], 22);
    const double u18_29 = bc(a[29], 22);
    const double u18_30 = bc(a[30], 22);
    if ((lm >> 0) & 1u) {
      const double m = a[0] * inv0;
      a[26] = a[26] - m * u0_26;
      a[0] = m;
    }
    if ((lm >> 1) & 1u) {
      const double m = a[1] * inv1;
      a[2] = a[2] - m * u1_2;
      a[24] = a[24] - m * u1_24;
      a[1] = m;
    }
    if ((lm >> 3) & 1u) {
      const double m = a[3] * inv3;
      a[4] = a[4] - m * u3_4;
      a[24] = a[24] - m * u3_24;
      a[3] = m;
    }
    if ((lm >> 5) & 1u) {
      const double m = a[5] * inv5;
      a[6] = a[6] - m * u5_6;
      a[24] = a[24] - m * u5_24;
      a[5] = m;
    }
    if ((lm >> 7) & 1u) {
      const double m = a[7] * inv7;
      a[24] = a[24] - m * u7_24;
      a[25] = a[25] - m * u7_25;
      a[7] = m;
    }
    if ((lm >> 8) & 1u) {
      const double m = a[8] * inv8;
      a[9] = a[9] - m * u8_9;
      a[24] = a[24] - m * u8_24;
      a[28] = a[28] - m * u8_28;
      a[8] = m;
    }
    if ((lm >> 10) & 1u) {
      const double m = a[10] * inv10;
      a[13] = a[13] - m * u10_13;
      a[25] = a[25] - m * u10_25;
      a[29] = a[29] - m * u10_29;
      a[10] = m;
    }
    if ((lm >> 11) & 1u) {
      const double m = a[11] * inv11;
      a[17] = a[17] - m * u11_17;
      a[24] = a[24] - m * u11_24;
      a[26] = a[26] - m * u11_26;
      a[11] = m;
    }
    if ((lm >> 12) & 1u) {
      const double m = a[12] * inv12;
      a[14] = a[14] - m * u12_14;
      a[24] = a[24] - m * u12_24;
      a[26] = a[26] - m * u12_26;
      a[12] = m;
    }
    if ((lm >> 15) & 1u) {
      const double m = a[15] * inv15;
      a[22] = a[22] - m * u15_22;
      a[23] = a[23] - m * u15_23;
      a[27] = a[27] - m * u15_27;
      a[29] = a[29] - m * u15_29;
      a[15] = m;
    }
    if ((lm >> 18) & 1u) {
      const double m = a[18] * inv18;
      a[19] = a[19] - m * u18_19;
      a[20] = a[20] - m * u18_20;
      a[21] = a[21] - m * u18_21;
      a[22] = a[22] - m * u18_22;
      a[24] = a[24] - m * u18_24;
      a[26] = a[26] - m * u18_26;
      a[28] = a[28] - m * u18_28;
      a[29] = a[29] - m * u18_29;
      a[30] = a[30] - m * u18_30;
      a[18] = m;
    }
  }
  {  // level 2: S+, Si+, Fe+, He+, H+, H2O+, CH+
    if (lvk == 2) a[pos] = 1.0 / a[pos];
    const double inv2 = bc(a[2], 24);
    const double u2_24 = bc(a[24], 24);
    const double inv4 = bc(a[4], 26);
    const double u4_24 = bc(a[24], 26);
    const double inv6 = bc(a[6], 28);
    const double u6_24 = bc(a[24], 28);
    const double inv9 = bc(a[9], 30);
    const double u9_24 = bc(a[24], 30);
    const double u9_28 = bc(a[28], 30);
    const double inv13 = bc(a[13], 3);
    const double u13_22 = bc(a[22], 3);
    const double u13_23 = bc(a[23], 3);
    const double u13_24 = bc(a[24], 3);
    const double u13_25 = bc(a[25], 3);
    const double u13_29 = bc(a[29], 3);
    const double u13_30 = bc(a[30], 3);
    const double inv14 = bc(a[14], 18);
    const double u14_16 = bc(a[16], 18);
    const double u14_24 = bc(a[24], 18);
    const double u14_26 = bc(a[26], 18);
    const double inv17 = bc(a[17], 11);
    const double u17_20 = bc(a[20], 11);
    const double u17_22 = bc(a[22], 11);
    const double u17_24 = bc(a[24], 11);
    const double u17_26 = bc(a[26], 11);
    const double u17_27 = bc(a[27], 11);
    if ((lm >> 2) & 1u) {
      const double m = a[2] * inv2;
      a[24] = a[24] - m * u2_24;
      a[2] = m;
    }
    if ((lm >> 4) & 1u) {
      const double m = a[4] * inv4;
      a[24] = a[24] - m * u4_24;
      a[4] = m;
    }
    if ((lm >> 6) & 1u) {
      const double m = a[6] * inv6;
      a[24] = a[24] - m * u6_24;
      a[6] = m;
    }
    if ((lm >> 9) & 1u) {
      const double m = a[9] * inv9;
      a[24] = a[24] - m * u9_24;
      a[28] = a[28] - m * u9_28;
      a[9] = m;
    }
    if ((lm >> 13) & 1u) {
      const double m = a[13] * inv13;
      a[22] = a[22] - m * u13_22;
      a[23] = a[23] - m * u13_23;
      a[24] = a[24] - m * u13_24;
      a[25] = a[25] - m * u13_25;
      a[29] = a[29] - m * u13_29;
      a[30] = a[30] - m * u13_30;
      a[13] = m;
    }
    if ((lm >> 14) & 1u) {
      const double m = a[14] * inv14;
      a[16] = a[16] - m * u14_16;
      a[24] = a[24] - m * u14_24;
      a[26] = a[26] - m * u14_26;
      a[14] = m;
    }
    if ((lm >> 17) & 1u) {
      const double m = a[17] * inv17;
      a[20] = a[20] - m * u17_20;
      a[22] = a[22] - m * u17_22;
      a[24] = a[24] - m * u17_24;
      a[26] = a[26] - m * u17_26;
      a[27] = a[27] - m * u17_27;
      a[17] = m;
    }
  }
  {  // level 3: OH+, CH
    if (lvk == 3) a[pos] = 1.0 / a[pos];
    const double inv16 = bc(a[16], 16);
    const double u16_20 = bc(a[20], 16);
    const double u16_24 = bc(a[24], 16);
    const double u16_26 = bc(a[26], 16);
    const double u16_29 = bc(a[29], 16);
    const double inv19 = bc(a[19], 10);
    const double u19_20 = bc(a[20], 10);
    const double u19_22 = bc(a[22], 10);
    const double u19_24 = bc(a[24], 10);
    const double u19_25 = bc(a[25], 10);
    const double u19_26 = bc(a[26], 10);
    const double u19_27 = bc(a[27], 10);
    const double u19_29 = bc(a[29], 10);
    if ((lm >> 16) & 1u) {
      const double m = a[16] * inv16;
      a[20] = a[20] - m * u16_20;
      a[24] = a[24] - m * u16_24;
      a[26] = a[26] - m * u16_26;
      a[29] = a[29] - m * u16_29;
      a[16] = m;
    }
    if ((lm >> 19) & 1u) {
      const double m = a[19] * inv19;
      a[20] = a[20] - m * u19_20;
      a[22] = a[22] - m * u19_22;
      a[24] = a[24] - m * u19_24;
      a[25] = a[25] - m * u19_25;
      a[26] = a[26] - m * u19_26;
      a[27] = a[27] - m * u19_27;
      a[29] = a[29] - m * u19_29;
      a[19] = m;
    }
  }
  {  // level 4: H3+
    if (lvk == 4) a[pos] = 1.0 / a[pos];
    const double inv20 = bc(a[20], 6);
    const double u20_24 = bc(a[24], 6);
    const double u20_26 = bc(a[26], 6);
    const double u20_27 = bc(a[27], 6);
    const double u20_28 = bc(a[28], 6);
    const double u20_29 = bc(a[29], 6);
    if ((lm >> 20) & 1u) {
      const double m = a[20] * inv20;
      a[24] = a[24] - m * u20_24;
      a[26] = a[26] - m * u20_26;
      a[27] = a[27] - m * u20_27;
      a[28] = a[28] - m * u20_28;
      a[29] = a[29] - m * u20_29;
      a[20] = m;
    }
  }
  {  // level 5: H2O
    if (lvk == 5) a[pos] = 1.0 / a[pos];
    const double inv21 = bc(a[21], 17);
    const double u21_22 = bc(a[22], 17);
    const double u21_23 = bc(a[23], 17);
    const double u21_24 = bc(a[24], 17);
    const double u21_26 = bc(a[26], 17);
    const double u21_27 = bc(a[27], 17);
    const double u21_28 = bc(a[28], 17);
    const double u21_29 = bc(a[29], 17);
    if ((lm >> 21) & 1u) {
      const double m = a[21] * inv21;
      a[22] = a[22] - m * u21_22;
      a[23] = a[23] - m * u21_23;
      a[24] = a[24] - m * u21_24;
      a[26] = a[26] - m * u21_26;
      a[27] = a[27] - m * u21_27;
      a[28] = a[28] - m * u21_28;
      a[29] = a[29] - m * u21_29;
      a[21] = m;
    }
  }
  {  // level 6: C+
    if (lvk == 6) a[pos] = 1.0 / a[pos];
    const double inv22 = bc(a[22], 7);
    const double u22_23 = bc(a[23], 7);
    const double u22_24 = bc(a[24], 7);
    const double u22_26 = bc(a[26], 7);
    const double u22_27 = bc(a[27], 7);
    const double u22_28 = bc(a[28], 7);
    const double u22_29 = bc(a[29], 7);
    if ((lm >> 22) & 1u) {
      const double m = a[22] * inv22;
      a[23] = a[23] - m * u22_23;
      a[24] = a[24] - m * u22_24;
      a[26] = a[26] - m * u22_26;
      a[27] = a[27] - m * u22_27;
      a[28] = a[28] - m * u22_28;
      a[29] = a[29] - m * u22_29;
      a[22] = m;
    }
  }
  {  // level 7: OH
    if (lvk == 7) a[pos] = 1.0 / a[pos];
    const double inv23 = bc(a[23], 15);
    const double u23_24 = bc(a[24], 15);
    const double u23_25 = bc(a[25], 15);
    const double u23_26 = bc(a[26], 15);
    const double u23_27 = bc(a[27], 15);
    const double u23_28 = bc(a[28], 15);
    const double u23_29 = bc(a[29], 15);
    if ((lm >> 23) & 1u) {
      const double m = a[23] * inv23;
      a[24] = a[24] - m * u23_24;
      a[25] = a[25] - m * u23_25;
      a[26] = a[26] - m * u23_26;
      a[27] = a[27] - m * u23_27;
      a[28] = a[28] - m * u23_28;
      a[29] = a[29] - m * u23_29;
      a[23] = m;
    }
  }
  {  // level 8: e-
    if (lvk == 8) a[pos] = 1.0 / a[pos];
    const double inv24 = bc(a[24], 0);
    const double u24_25 = bc(a[25], 0);
    const double u24_26 = bc(a[26], 0);
    const double u24_27 = bc(a[27], 0);
    const double u24_28 = bc(a[28], 0);
    const double u24_29 = bc(a[29], 0);
    const double u24_30 = bc(a[30], 0);
    if ((lm >> 24) & 1u) {
      const double m = a[24] * inv24;
      a[25] = a[25] - m * u24_25;
      a[26] = a[26] - m * u24_26;
      a[27] = a[27] - m * u24_27;
      a[28] = a[28] - m * u24_28;
      a[29] = a[29] - m * u24_29;
      a[30] = a[30] - m * u24_30;
      a[24] = m;
    }
  }
  {  // level 9: H
    if (lvk == 9) a[pos] = 1.0 / a[pos];
    const double inv25 = bc(a[25], 1);
    const double u25_26 = bc(a[26], 1);
    const double u25_27 = bc(a[27], 1);
    const double u25_28 = bc(a[28], 1);
    const double u25_29 = bc(a[29], 1);
    const double u25_30 = bc(a[30], 1);
    if ((lm >> 25) & 1u) {
      const double m = a[25] * inv25;
      a[26] = a[26] - m * u25_26;
      a[27] = a[27] - m * u25_27;
      a[28] = a[28] - m * u25_28;
      a[29] = a[29] - m * u25_29;
      a[30] = a[30] - m * u25_30;
      a[25] = m;
    }
  }
  {  // level 10: H2
    if (lvk == 10) a[pos] = 1.0 / a[pos];
    const double inv26 = bc(a[26], 2);
    const double u26_27 = bc(a[27], 2);
    const double u26_28 = bc(a[28], 2);
    const double u26_29 = bc(a[29], 2);
    const double u26_30 = bc(a[30], 2);
    if ((lm >> 26) & 1u) {
      const double m = a[26] * inv26;
      a[27] = a[27] - m * u26_27;
      a[28] = a[28] - m * u26_28;
      a[29] = a[29] - m * u26_29;
      a[30] = a[30] - m * u26_30;
      a[26] = m;
    }
  }
  {  // level 11: C
    if (lvk == 11) a[pos] = 1.0 / a[pos];
    const double inv27 = bc(a[27], 8);
    const double u27_28 = bc(a[28], 8);
    const double u27_29 = bc(a[29], 8);
    const double u27_30 = bc(a[30], 8);
    if ((lm >> 27) & 1u) {
      const double m = a[27] * inv27;
      a[28] = a[28] - m * u27_28;
      a[29] = a[29] - m * u27_29;
      a[30] = a[30] - m * u27_30;
      a[27] = m;
    }
  }
  {  // level 12: CO
    if (lvk == 12) a[pos] = 1.0 / a[pos];
    const double inv28 = bc(a[28], 9);
    const double u28_29 = bc(a[29], 9);
    const double u28_30 = bc(a[30], 9);
    if ((lm >> 28) & 1u) {
      const double m = a[28] * inv28;
      a[29] = a[29] - m * u28_29;
      a[30] = a[30] - m * u28_30;
      a[28] = m;
    }
  }
  {  // level 13: O
    if (lvk == 13) a[pos] = 1.0 / a[pos];
    const double inv29 = bc(a[29], 13);
    const double u29_30 = bc(a[30], 13);
    if ((lm >> 29) & 1u) {
      const double m = a[29] * inv29;
      a[30] = a[30] - m * u29_30;
      a[29] = m;
    }
  }
  {  // level 14: CO+
    if (lvk == 14) a[pos] = 1.0 / a[pos];
    const double inv30 = bc(a[30], 21);
  }
}

// per lane: elimination position (31: no row) and the forward | backward << 8
// dependency levels of the solves | the factorisation's level << 16
#define P3D_WAVE_POS {24, 25, 26, 13, 7, 0, 20, 22, 27, 28, 19, 17, 11, 29, 10, 23, 16, 21, 14, 12, 15, 30, 18, 1, 2, 3, 4, 5, 6, 8, 9, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31, 31}
#define P3D_WAVE_LEVELS {526086, 591367, 656648, 133634, 67585, 67073, 264194, 395524, 721929, 787210, 199170, 133633, 68353, 852491, 68353, 460805, 198913, 330243, 133633, 68353, 68097, 917772, 68353, 67841, 133122, 67841, 133122, 67841, 133122, 67841, 133122, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}

// Solve M x = b with the factors; b and x: the lane's species component.
// Level by level: the components of one level are broadcast, each lane of
// the level records its component in xs[pos], then every lane subtracts
// a[k] x_k with no mask.  A lane reached by a component it has no entry
// for in this triangle (an entry of the other triangle, or its own
// diagonal) is always final by then, so only its register is spoilt and
// the recorded xs stays exact; the few components for which that is not
// so keep their mask.  dinv: the lane's 1 / u_ii.
template <class A, class BC>
__device__ __forceinline__ double lu_solve(const A &a, unsigned lm, unsigned um, int lane,
                                          double b, const BC &bc, double *xs, int pos,
                                          int lev, double dinv) {
  const int lvf = lev & 0xff, lvb = (lev >> 8) & 0xff;
  {
    const double x0 = bc(b, 5);
    const double x1 = bc(b, 23);
    const double x3 = bc(b, 25);
    const double x5 = bc(b, 27);
    const double x7 = bc(b, 4);
    const double x8 = bc(b, 29);
    const double x10 = bc(b, 14);
    const double x11 = bc(b, 12);
    const double x12 = bc(b, 19);
    const double x14 = bc(b, 18);
    const double x15 = bc(b, 20);
    const double x16 = bc(b, 16);
    const double x17 = bc(b, 11);
    const double x18 = bc(b, 22);
    if (lvf == 1) xs[pos] = b;
    b -= a[0] * x0;
    b -= a[1] * x1;
    b -= a[3] * x3;
    b -= a[5] * x5;
    b -= a[7] * x7;
    b -= a[8] * x8;
    b -= a[10] * x10;
    b -= a[11] * x11;
    b -= a[12] * x12;
    b -= a[14] * x14;
    b -= a[15] * x15;
    b -= a[16] * x16;
    b -= a[17] * x17;
    b -= a[18] * x18;
  }
  {
    const double x2 = bc(b, 24);
    const double x4 = bc(b, 26);
    const double x6 = bc(b, 28);
    const double x9 = bc(b, 30);
    const double x13 = bc(b, 3);
    const double x19 = bc(b, 10);
    const double x20 = bc(b, 6);
    if (lvf == 2) xs[pos] = b;
    b -= a[2] * x2;
    b -= a[4] * x4;
    b -= a[6] * x6;
    b -= a[9] * x9;
    b -= a[13] * x13;
    b -= a[19] * x19;
    b -= a[20] * x20;
  }
  {
    const double x21 = bc(b, 17);
    if (lvf == 3) xs[pos] = b;
    b -= a[21] * x21;
  }
  {
    const double x22 = bc(b, 7);
    if (lvf == 4) xs[pos] = b;
    b -= a[22] * x22;
  }
  {
    const double x23 = bc(b, 15);
    if (lvf == 5) xs[pos] = b;
    b -= a[23] * x23;
  }
  {
    const double x24 = bc(b, 0);
    if (lvf == 6) xs[pos] = b;
    b -= a[24] * x24;
  }
  {
    const double x25 = bc(b, 1);
    if (lvf == 7) xs[pos] = b;
    b -= a[25] * x25;
  }
  {
    const double x26 = bc(b, 2);
    if (lvf == 8) xs[pos] = b;
    b -= a[26] * x26;
  }
  {
    const double x27 = bc(b, 8);
    if (lvf == 9) xs[pos] = b;
    b -= a[27] * x27;
  }
  {
    const double x28 = bc(b, 9);
    if (lvf == 10) xs[pos] = b;
    b -= a[28] * x28;
  }
  {
    const double x29 = bc(b, 13);
    if (lvf == 11) xs[pos] = b;
    b -= a[29] * x29;
  }
  {
    if (lvf == 12) xs[pos] = b;
  }
  b = xs[pos];
  {
    const double bx = b * dinv;
    if (lvb == 1) xs[pos] = bx;
    const double x30 = bc(bx, 21);
    b -= a[30] * x30;
  }
  {
    const double bx = b * dinv;
    if (lvb == 2) xs[pos] = bx;
    const double x29 = bc(bx, 13);
    b -= a[29] * x29;
  }
  {
    const double bx = b * dinv;
    if (lvb == 3) xs[pos] = bx;
    const double x28 = bc(bx, 9);
    b -= a[28] * x28;
  }
  {
    const double bx = b * dinv;
    if (lvb == 4) xs[pos] = bx;
    const double x27 = bc(bx, 8);
    b -= a[27] * x27;
  }
  {
    const double bx = b * dinv;
    if (lvb == 5) xs[pos] = bx;
    const double x26 = bc(bx, 2);
    b -= a[26] * x26;
  }
  {
    const double bx = b * dinv;
    if (lvb == 6) xs[pos] = bx;
    const double x25 = bc(bx, 1);
    b -= a[25] * x25;
  }
  {
    const double bx = b * dinv;
    if (lvb == 7) xs[pos] = bx;
    const double x24 = bc(bx, 0);
    b -= a[24] * x24;
  }
  {
    const double bx = b * dinv;
    if (lvb == 8) xs[pos] = bx;
    const double x23 = bc(bx, 15);
    const double x20 = bc(bx, 6);
    const double x9 = bc(bx, 30);
    const double x6 = bc(bx, 28);
    const double x4 = bc(bx, 26);
    const double x2 = bc(bx, 24);
    b -= a[23] * x23;
    if ((um >> 20) & 1u) b -= a[20] * x20;
    if ((um >> 9) & 1u) b -= a[9] * x9;
    b -= a[6] * x6;
    b -= a[4] * x4;
    b -= a[2] * x2;
  }
  {
    const double bx = b * dinv;
    if (lvb == 9) xs[pos] = bx;
    const double x22 = bc(bx, 7);
    const double x16 = bc(bx, 16);
    b -= a[22] * x22;
    if ((um >> 16) & 1u) b -= a[16] * x16;
  }
  {
    const double bx = b * dinv;
    if (lvb == 10) xs[pos] = bx;
    const double x21 = bc(bx, 17);
    const double x19 = bc(bx, 10);
    const double x17 = bc(bx, 11);
    const double x14 = bc(bx, 18);
    const double x13 = bc(bx, 3);
    b -= a[21] * x21;
    b -= a[19] * x19;
    b -= a[17] * x17;
    b -= a[14] * x14;
    b -= a[13] * x13;
  }
  {
    const double bx = b * dinv;
    if (lvb == 11) xs[pos] = bx;
  }
  return xs[pos];
}

// Transposed solve: per species (lane), bit i where its column of the
// factors holds a U^T (upper of the lane's row of U^T) or L^T entry, and
// the forward | backward << 8 levels of the transposed solves
#define P3D_WAVE_TUMASK {0x00ff7bfeu, 0x01882480u, 0x03ff5801u, 0x00000400u, 0x00000000u, 0x00000000u, 0x000f0000u, 0x002ea000u, 0x07fa8000u, 0x0ff40300u, 0x00040000u, 0x00000800u, 0x00000000u, 0x1ffda400u, 0x00000000u, 0x0060a000u, 0x00004000u, 0x00040000u, 0x00001000u, 0x00000000u, 0x00000000u, 0x3f042000u, 0x00000000u, 0x00000000u, 0x00000002u, 0x00000000u, 0x00000008u, 0x00000000u, 0x00000020u, 0x00000000u, 0x00000100u}
#define P3D_WAVE_TLMASK {0x7e000000u, 0x7c000000u, 0x78000000u, 0x23000000u, 0x07000000u, 0x06100000u, 0x7fa00000u, 0x7f800000u, 0x70000000u, 0x60000000u, 0x7f000000u, 0x0f080000u, 0x0b080000u, 0x40000000u, 0x22002000u, 0x7f000000u, 0x27a00000u, 0x5bc00000u, 0x27a00000u, 0x03a00000u, 0x78400000u, 0x00000000u, 0x53000000u, 0x01000004u, 0x01000000u, 0x01000010u, 0x01000000u, 0x01000040u, 0x01000000u, 0x01000200u, 0x31400000u}
#define P3D_WAVE_TLEVELS {1797, 1542, 1287, 2050, 2049, 3073, 2820, 2307, 1032, 777, 2050, 2306, 2305, 522, 2305, 2052, 2819, 2562, 2818, 2817, 2561, 267, 2049, 2305, 2050, 2305, 2050, 2305, 2050, 2817, 2562, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}

// Solve M^T x = b with the factors of lu_factor in the column layout: a[i]
// is entry (i, pos) of the factored matrix in elimination positions, i.e.
// u_(i,pos) above the diagonal, l_(i,pos) below it and 1 / u_(pos,pos) on
// it (dinv).  U^T z = b runs level by level with the division, then
// L^T x = z; components are recorded in xs as in lu_solve, and the
// components that could reach a lane that is not final keep a mask (um:
// the lane's U^T entries, lm: its L^T entries).
template <class BC>
__device__ __forceinline__ double lu_solve_t(const double (&a)[32], unsigned lm, unsigned um,
                                            int lane, double b, const BC &bc, double *xs,
                                            int pos, int lev, double dinv) {
  const int lvf = lev & 0xff, lvb = lev >> 8;
  {
    const double bx = b * dinv;
    if (lvf == 1) xs[pos] = bx;
    const double x0 = bc(bx, 5);
    const double x1 = bc(bx, 23);
    const double x3 = bc(bx, 25);
    const double x5 = bc(bx, 27);
    const double x7 = bc(bx, 4);
    const double x8 = bc(bx, 29);
    const double x10 = bc(bx, 14);
    const double x11 = bc(bx, 12);
    const double x12 = bc(bx, 19);
    const double x15 = bc(bx, 20);
    const double x18 = bc(bx, 22);
    b -= a[0] * x0;
    b -= a[1] * x1;
    b -= a[3] * x3;
    b -= a[5] * x5;
    b -= a[7] * x7;
    b -= a[8] * x8;
    b -= a[10] * x10;
    b -= a[11] * x11;
    b -= a[12] * x12;
    b -= a[15] * x15;
    b -= a[18] * x18;
  }
  {
    const double bx = b * dinv;
    if (lvf == 2) xs[pos] = bx;
    const double x2 = bc(bx, 24);
    const double x4 = bc(bx, 26);
    const double x6 = bc(bx, 28);
    const double x9 = bc(bx, 30);
    const double x13 = bc(bx, 3);
    const double x14 = bc(bx, 18);
    const double x17 = bc(bx, 11);
    const double x19 = bc(bx, 10);
    const double x21 = bc(bx, 17);
    b -= a[2] * x2;
    b -= a[4] * x4;
    b -= a[6] * x6;
    b -= a[9] * x9;
    b -= a[13] * x13;
    b -= a[14] * x14;
    b -= a[17] * x17;
    b -= a[19] * x19;
    if ((um >> 21) & 1u) b -= a[21] * x21;
  }
  {
    const double bx = b * dinv;
    if (lvf == 3) xs[pos] = bx;
    const double x16 = bc(bx, 16);
    const double x22 = bc(bx, 7);
    b -= a[16] * x16;
    b -= a[22] * x22;
  }
  {
    const double bx = b * dinv;
    if (lvf == 4) xs[pos] = bx;
    const double x20 = bc(bx, 6);
    const double x23 = bc(bx, 15);
    b -= a[20] * x20;
    b -= a[23] * x23;
  }
  {
    const double bx = b * dinv;
    if (lvf == 5) xs[pos] = bx;
    const double x24 = bc(bx, 0);
    b -= a[24] * x24;
  }
  {
    const double bx = b * dinv;
    if (lvf == 6) xs[pos] = bx;
    const double x25 = bc(bx, 1);
    b -= a[25] * x25;
  }
  {
    const double bx = b * dinv;
    if (lvf == 7) xs[pos] = bx;
    const double x26 = bc(bx, 2);
    b -= a[26] * x26;
  }
  {
    const double bx = b * dinv;
    if (lvf == 8) xs[pos] = bx;
    const double x27 = bc(bx, 8);
    b -= a[27] * x27;
  }
  {
    const double bx = b * dinv;
    if (lvf == 9) xs[pos] = bx;
    const double x28 = bc(bx, 9);
    b -= a[28] * x28;
  }
  {
    const double bx = b * dinv;
    if (lvf == 10) xs[pos] = bx;
    const double x29 = bc(bx, 13);
    b -= a[29] * x29;
  }
  {
    const double bx = b * dinv;
    if (lvf == 11) xs[pos] = bx;
  }
  b = xs[pos];
  {
    if (lvb == 1) xs[pos] = b;
    const double x30 = bc(b, 21);
    b -= a[30] * x30;
  }
  {
    if (lvb == 2) xs[pos] = b;
    const double x29 = bc(b, 13);
    b -= a[29] * x29;
  }
  {
    if (lvb == 3) xs[pos] = b;
    const double x28 = bc(b, 9);
    b -= a[28] * x28;
  }
  {
    if (lvb == 4) xs[pos] = b;
    const double x27 = bc(b, 8);
    b -= a[27] * x27;
  }
  {
    if (lvb == 5) xs[pos] = b;
    const double x26 = bc(b, 2);
    b -= a[26] * x26;
  }
  {
    if (lvb == 6) xs[pos] = b;
    const double x25 = bc(b, 1);
    b -= a[25] * x25;
  }
  {
    if (lvb == 7) xs[pos] = b;
    const double x24 = bc(b, 0);
    b -= a[24] * x24;
  }
  {
    if (lvb == 8) xs[pos] = b;
    const double x23 = bc(b, 15);
    const double x19 = bc(b, 10);
    const double x13 = bc(b, 3);
    const double x6 = bc(b, 28);
    const double x4 = bc(b, 26);
    const double x2 = bc(b, 24);
    b -= a[23] * x23;
    if ((lm >> 19) & 1u) b -= a[19] * x19;
    if ((lm >> 13) & 1u) b -= a[13] * x13;
    b -= a[6] * x6;
    b -= a[4] * x4;
    b -= a[2] * x2;
  }
  {
    if (lvb == 9) xs[pos] = b;
    const double x22 = bc(b, 7);
    b -= a[22] * x22;
  }
  {
    if (lvb == 10) xs[pos] = b;
    const double x21 = bc(b, 17);
    const double x9 = bc(b, 30);
    b -= a[21] * x21;
    b -= a[9] * x9;
  }
  {
    if (lvb == 11) xs[pos] = b;
    const double x20 = bc(b, 6);
    b -= a[20] * x20;
  }
  {
    if (lvb == 12) xs[pos] = b;
  }
  return xs[pos];
}

}  // namespace wave
}  // namespace p3d
