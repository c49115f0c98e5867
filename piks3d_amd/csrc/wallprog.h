// wallprog.h -- cell labels, octant signs, in-sweep wall switches and the
// diffuse/specular wall boundary programs of the DVM iteration, as constexpr
// tables usable from host and device code.
//
// Reference: labels src/physicalGrid.F90:127-134; octant sign conventions
// src/solver.F90:174-485 (comment headers); switches src/solver.F90:184-513;
// wall boundary condition src/solver.F90:761-1826.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define PIKS_HD __host__ __device__
#else
#define PIKS_HD
#endif

namespace piks {

enum Label : int {
  FLUID = 0, SOLID = 1, GHOST = 4,
  L_E = 20, L_W = 21, L_N = 22, L_S = 23, L_F = 24, L_B = 25,
  L_EN = 30, L_WN = 31, L_ES = 32, L_WS = 33,
  L_NF = 40, L_NB = 41, L_SF = 42, L_SB = 43,
  L_EF = 50, L_EB = 51, L_WF = 52, L_WB = 53,
  L_ENF = 60, L_WNF = 61, L_WSF = 62, L_ESF = 63,
  L_ENB = 70, L_WNB = 71, L_WSB = 72, L_ESB = 73
};

// Octants 0..7 (the reference's 1..8): signs of (cx, cy, cz).
PIKS_HD constexpr int osx(int o) { return (o == 1 || o == 2 || o == 5 || o == 6) ? -1 : 1; }
PIKS_HD constexpr int osy(int o) { return (o == 2 || o == 3 || o == 6 || o == 7) ? -1 : 1; }
PIKS_HD constexpr int osz(int o) { return o >= 4 ? -1 : 1; }
// Octants whose velocity signs are equal or opposite share a storage order
// (family 0: octants 0, 6; 1: 1, 7; 2: 2, 4; 3: 3, 5); the family's signs are
// those of octant `family`.
PIKS_HD constexpr int ofam(int o) { return o < 4 ? o : (o == 4 ? 2 : o == 5 ? 3 : o == 6 ? 0 : 1); }

PIKS_HD constexpr bool is_wall(int lab) { return lab >= 20; }
PIKS_HD constexpr bool is_corner3(int lab) { return lab >= 60; }

// Which fluid sides a wall label has (bit 0 E, 1 W, 2 N, 3 S, 4 F, 5 B).
PIKS_HD constexpr int wall_sides(int lab) {
  switch (lab) {
    case L_E: return 1; case L_W: return 2; case L_N: return 4; case L_S: return 8;
    case L_F: return 16; case L_B: return 32;
    case L_EN: return 1 | 4; case L_WN: return 2 | 4; case L_ES: return 1 | 8; case L_WS: return 2 | 8;
    case L_NF: return 4 | 16; case L_NB: return 4 | 32; case L_SF: return 8 | 16; case L_SB: return 8 | 32;
    case L_EF: return 1 | 16; case L_EB: return 1 | 32; case L_WF: return 2 | 16; case L_WB: return 2 | 32;
    case L_ENF: return 1 | 4 | 16; case L_WNF: return 2 | 4 | 16; case L_WSF: return 2 | 8 | 16;
    case L_ESF: return 1 | 8 | 16; case L_ENB: return 1 | 4 | 32; case L_WNB: return 2 | 4 | 32;
    case L_WSB: return 2 | 8 | 32; case L_ESB: return 1 | 8 | 32;
    default: return 0;
  }
}

// Label from the neighbour product 2^E 3^W 5^N 7^S 9^F 11^B
// (src/physicalGrid.F90:554-667); -1 for products the reference leaves solid.
PIKS_HD constexpr int label_of_product(int p) {
  switch (p) {
    case 2: return L_E; case 3: return L_W; case 5: return L_N; case 7: return L_S;
    case 9: return L_F; case 11: return L_B;
    case 10: return L_EN; case 15: return L_WN; case 14: return L_ES; case 21: return L_WS;
    case 18: return L_EF; case 22: return L_EB; case 27: return L_WF; case 33: return L_WB;
    case 45: return L_NF; case 55: return L_NB; case 63: return L_SF; case 77: return L_SB;
    case 90: return L_ENF; case 110: return L_ENB; case 135: return L_WNF; case 165: return L_WNB;
    case 126: return L_ESF; case 154: return L_ESB; case 189: return L_WSF; case 231: return L_WSB;
    default: return -1;
  }
}

// In-sweep switch for octant o reading wall label `lab` along y (axis 1) or z
// (axis 2): returns the source slot 1..8 of the same wall cell, 9 for the
// 3-fold corner store fw(corner, o), 0 if the reference does not switch.
PIKS_HD constexpr int switch_source(int o, int axis, int lab) {
  if (axis == 1) {
    switch (o) {
      case 0: return lab == L_EN ? 3 : (lab == L_ENF || lab == L_ENB) ? 9 : 0;
      case 1: return lab == L_WN ? 4 : (lab == L_WNF || lab == L_WNB) ? 9 : 0;
      case 2: return lab == L_WS ? 1 : (lab == L_WSF || lab == L_WSB) ? 9 : 0;
      case 3: return lab == L_ES ? 2 : (lab == L_ESF || lab == L_ESB) ? 9 : 0;
      case 4: return lab == L_EN ? 7 : (lab == L_ENF || lab == L_ENB) ? 9 : 0;
      case 5: return lab == L_WN ? 8 : (lab == L_WNF || lab == L_WNB) ? 9 : 0;
      case 6: return lab == L_WS ? 8 : (lab == L_WSF || lab == L_WSB) ? 9 : 0;
      case 7: return lab == L_ES ? 6 : (lab == L_ESF || lab == L_ESB) ? 9 : 0;
    }
  } else if (axis == 2) {
    switch (o) {
      case 0: return lab == L_NF ? 8 : lab == L_EF ? 6 : (lab == L_WNF || lab == L_ESF) ? 9 : lab == L_ENF ? 7 : 0;
      case 1: return lab == L_NF ? 7 : lab == L_WF ? 5 : (lab == L_ENF || lab == L_WSF) ? 9 : lab == L_WNF ? 8 : 0;
      case 2: return lab == L_SF ? 6 : lab == L_WF ? 8 : (lab == L_WNF || lab == L_ESF) ? 9 : lab == L_WSF ? 5 : 0;
      case 3: return lab == L_SF ? 5 : lab == L_EF ? 7 : (lab == L_ENF || lab == L_WSF) ? 9 : lab == L_ESF ? 6 : 0;
      case 4: return lab == L_NB ? 4 : lab == L_EB ? 2 : (lab == L_ESB || lab == L_WNB) ? 9 : lab == L_ENB ? 3 : 0;
      case 5: return lab == L_NB ? 3 : lab == L_WB ? 1 : (lab == L_ENB || lab == L_WSB) ? 9 : lab == L_WNB ? 4 : 0;
      case 6: return lab == L_SB ? 2 : lab == L_WB ? 4 : (lab == L_WNB || lab == L_ESB) ? 9 : lab == L_WSB ? 1 : 0;
      case 7: return lab == L_SB ? 1 : lab == L_EB ? 3 : (lab == L_ENB || lab == L_WSB) ? 9 : lab == L_ESB ? 2 : 0;
    }
  }
  return 0;
}

// Wall boundary program.  Value slots: 0..7 f(W, o), 8..15 fw(corner, o),
// 16..23 a per-wall temporary.  Each face extrapolates its four incoming
// octants from the two fluid-side neighbours into `store`+octant and sums
// c_axis * (v0+v1+v2+v3) in that order into R[face]; the writes then run in
// order: dst = accom*w*R[rho] + (1-accom)*src.
struct Face { int8_t axis, dir, store, in[4]; };
struct Wr { int8_t dst, rho, src; };
struct WallProg { int8_t nface; Face face[3]; int8_t nwr; Wr wr[12]; };

#define PK_F(n) int8_t((n) - 1)
#define PK_W(n) int8_t(8 + (n) - 1)
#define PK_Z(n) int8_t(16 + (n) - 1)
#define PK_XP Face{0, 1, 0, {1, 2, 5, 6}}
#define PK_XM Face{0, -1, 0, {0, 3, 4, 7}}
#define PK_YP(s) Face{1, 1, s, {2, 3, 6, 7}}
#define PK_YM(s) Face{1, -1, s, {0, 1, 4, 5}}
#define PK_ZP(s) Face{2, 1, s, {4, 5, 6, 7}}
#define PK_ZM(s) Face{2, -1, s, {0, 1, 2, 3}}
#define PK_OUTXP Wr{PK_F(1), 0, PK_F(2)}, Wr{PK_F(4), 0, PK_F(3)}, Wr{PK_F(5), 0, PK_F(6)}, Wr{PK_F(8), 0, PK_F(7)}
#define PK_OUTXM Wr{PK_F(2), 0, PK_F(1)}, Wr{PK_F(3), 0, PK_F(4)}, Wr{PK_F(6), 0, PK_F(5)}, Wr{PK_F(7), 0, PK_F(8)}
#define PK_OUTYP Wr{PK_F(1), 0, PK_F(4)}, Wr{PK_F(2), 0, PK_F(3)}, Wr{PK_F(5), 0, PK_F(8)}, Wr{PK_F(6), 0, PK_F(7)}
#define PK_OUTYM Wr{PK_F(3), 0, PK_F(2)}, Wr{PK_F(4), 0, PK_F(1)}, Wr{PK_F(7), 0, PK_F(6)}, Wr{PK_F(8), 0, PK_F(5)}

PIKS_HD constexpr WallProg wall_prog(int lab) {
  switch (lab) {
    case L_E: return WallProg{1, {PK_XP}, 4, {PK_OUTXP}};
    case L_W: return WallProg{1, {PK_XM}, 4, {PK_OUTXM}};
    case L_N: return WallProg{1, {PK_YP(0)}, 4, {PK_OUTYP}};
    case L_S: return WallProg{1, {PK_YM(0)}, 4, {PK_OUTYM}};
    case L_F: return WallProg{1, {PK_ZP(0)}, 4, {Wr{PK_F(1), 0, PK_F(5)}, Wr{PK_F(2), 0, PK_F(6)}, Wr{PK_F(3), 0, PK_F(7)}, Wr{PK_F(4), 0, PK_F(8)}}};
    case L_B: return WallProg{1, {PK_ZM(0)}, 4, {Wr{PK_F(5), 0, PK_F(1)}, Wr{PK_F(6), 0, PK_F(2)}, Wr{PK_F(7), 0, PK_F(3)}, Wr{PK_F(8), 0, PK_F(4)}}};
    case L_EN: return WallProg{2, {PK_XP, PK_YP(16)}, 8, {PK_OUTXP,
        Wr{PK_F(2), 1, PK_Z(3)}, Wr{PK_F(6), 1, PK_Z(7)}, Wr{PK_F(3), 1, PK_Z(4)}, Wr{PK_F(7), 1, PK_Z(8)}}};
    case L_WN: return WallProg{2, {PK_XM, PK_YP(16)}, 8, {PK_OUTXM,
        Wr{PK_F(1), 1, PK_Z(4)}, Wr{PK_F(5), 1, PK_Z(8)}, Wr{PK_F(4), 1, PK_Z(3)}, Wr{PK_F(8), 1, PK_Z(7)}}};
    case L_ES: return WallProg{2, {PK_XP, PK_YM(16)}, 8, {PK_OUTXP,
        Wr{PK_F(3), 1, PK_Z(2)}, Wr{PK_F(7), 1, PK_Z(6)}, Wr{PK_F(2), 1, PK_Z(1)}, Wr{PK_F(6), 1, PK_Z(5)}}};
    case L_WS: return WallProg{2, {PK_XM, PK_YM(16)}, 8, {PK_OUTXM,
        Wr{PK_F(4), 1, PK_Z(1)}, Wr{PK_F(8), 1, PK_Z(5)}, Wr{PK_F(1), 1, PK_Z(2)}, Wr{PK_F(5), 1, PK_Z(6)}}};
    case L_NF: return WallProg{2, {PK_YP(0), PK_ZP(16)}, 8, {PK_OUTYP,
        Wr{PK_F(3), 1, PK_Z(7)}, Wr{PK_F(4), 1, PK_Z(8)}, Wr{PK_F(7), 1, PK_Z(6)}, Wr{PK_F(8), 1, PK_Z(5)}}};
    case L_NB: return WallProg{2, {PK_YP(0), PK_ZM(16)}, 8, {PK_OUTYP,
        Wr{PK_F(7), 1, PK_Z(3)}, Wr{PK_F(8), 1, PK_Z(4)}, Wr{PK_F(4), 1, PK_Z(1)}, Wr{PK_F(3), 1, PK_Z(2)}}};
    case L_SF: return WallProg{2, {PK_YM(0), PK_ZP(16)}, 8, {PK_OUTYM,
        Wr{PK_F(1), 1, PK_Z(5)}, Wr{PK_F(2), 1, PK_Z(6)}, Wr{PK_F(6), 1, PK_Z(7)}, Wr{PK_F(5), 1, PK_Z(8)}}};
    case L_SB: return WallProg{2, {PK_YM(0), PK_ZM(16)}, 8, {PK_OUTYM,
        Wr{PK_F(5), 1, PK_Z(1)}, Wr{PK_F(6), 1, PK_Z(2)}, Wr{PK_F(2), 1, PK_Z(3)}, Wr{PK_F(1), 1, PK_Z(4)}}};
    case L_EF: return WallProg{2, {PK_XP, PK_ZP(16)}, 8, {PK_OUTXP,
        Wr{PK_F(2), 1, PK_Z(6)}, Wr{PK_F(3), 1, PK_Z(7)}, Wr{PK_F(6), 1, PK_Z(5)}, Wr{PK_F(7), 1, PK_Z(8)}}};
    case L_EB: return WallProg{2, {PK_XP, PK_ZM(16)}, 8, {PK_OUTXP,
        Wr{PK_F(6), 1, PK_Z(2)}, Wr{PK_F(7), 1, PK_Z(3)}, Wr{PK_F(2), 1, PK_Z(1)}, Wr{PK_F(3), 1, PK_Z(4)}}};
    case L_WF: return WallProg{2, {PK_XM, PK_ZP(16)}, 8, {PK_OUTXM,
        Wr{PK_F(1), 1, PK_Z(5)}, Wr{PK_F(4), 1, PK_Z(8)}, Wr{PK_F(5), 1, PK_Z(6)}, Wr{PK_F(8), 1, PK_Z(7)}}};
    case L_WB: return WallProg{2, {PK_XM, PK_ZM(16)}, 8, {PK_OUTXM,
        Wr{PK_F(5), 1, PK_Z(1)}, Wr{PK_F(8), 1, PK_Z(4)}, Wr{PK_F(1), 1, PK_Z(2)}, Wr{PK_F(4), 1, PK_Z(3)}}};
    case L_ENF: return WallProg{3, {PK_XP, PK_YP(8), PK_ZP(16)}, 12, {PK_OUTXP,
        Wr{PK_F(2), 1, PK_W(3)}, Wr{PK_F(6), 1, PK_W(7)}, Wr{PK_F(3), 2, PK_Z(7)},
        Wr{PK_W(1), 1, PK_W(4)}, Wr{PK_W(5), 1, PK_W(8)}, Wr{PK_W(2), 2, PK_Z(6)}, Wr{PK_W(4), 2, PK_Z(8)},
        Wr{PK_F(7), 2, PK_Z(5)}}};
    case L_WNF: return WallProg{3, {PK_XM, PK_YP(8), PK_ZP(16)}, 12, {PK_OUTXM,
        Wr{PK_F(1), 1, PK_W(4)}, Wr{PK_F(5), 1, PK_W(8)}, Wr{PK_F(4), 2, PK_Z(8)},
        Wr{PK_W(2), 1, PK_W(3)}, Wr{PK_W(6), 1, PK_W(7)}, Wr{PK_W(1), 2, PK_Z(5)}, Wr{PK_W(3), 2, PK_Z(7)},
        Wr{PK_F(8), 2, PK_Z(6)}}};
    case L_WSF: return WallProg{3, {PK_XM, PK_YM(8), PK_ZP(16)}, 12, {PK_OUTXM,
        Wr{PK_F(4), 1, PK_W(1)}, Wr{PK_F(8), 1, PK_W(5)}, Wr{PK_F(1), 2, PK_Z(5)},
        Wr{PK_W(3), 1, PK_W(2)}, Wr{PK_W(7), 1, PK_W(6)}, Wr{PK_W(2), 2, PK_Z(6)}, Wr{PK_W(4), 2, PK_Z(8)},
        Wr{PK_F(5), 2, PK_Z(7)}}};
    case L_ESF: return WallProg{3, {PK_XP, PK_YM(8), PK_ZP(16)}, 12, {PK_OUTXP,
        Wr{PK_F(3), 1, PK_W(2)}, Wr{PK_F(7), 1, PK_W(6)}, Wr{PK_F(2), 2, PK_Z(6)},
        Wr{PK_W(4), 1, PK_W(1)}, Wr{PK_W(8), 1, PK_W(5)}, Wr{PK_W(1), 2, PK_Z(5)}, Wr{PK_W(3), 2, PK_Z(7)},
        Wr{PK_F(6), 2, PK_Z(8)}}};
    case L_ENB: return WallProg{3, {PK_XP, PK_YP(8), PK_ZM(16)}, 12, {PK_OUTXP,
        Wr{PK_F(2), 1, PK_W(3)}, Wr{PK_F(6), 1, PK_W(7)}, Wr{PK_F(7), 2, PK_Z(3)},
        Wr{PK_W(1), 1, PK_W(4)}, Wr{PK_W(5), 1, PK_W(8)}, Wr{PK_W(6), 2, PK_Z(2)}, Wr{PK_W(8), 2, PK_Z(4)},
        Wr{PK_F(3), 2, PK_Z(1)}}};
    case L_WNB: return WallProg{3, {PK_XM, PK_YP(8), PK_ZM(16)}, 12, {PK_OUTXM,
        Wr{PK_F(1), 1, PK_W(4)}, Wr{PK_F(5), 1, PK_W(8)}, Wr{PK_F(8), 2, PK_Z(4)},
        Wr{PK_W(2), 1, PK_W(3)}, Wr{PK_W(6), 1, PK_W(7)}, Wr{PK_W(5), 2, PK_Z(1)}, Wr{PK_W(7), 2, PK_Z(3)},
        Wr{PK_F(4), 2, PK_Z(2)}}};
    case L_WSB: return WallProg{3, {PK_XM, PK_YM(8), PK_ZM(16)}, 12, {PK_OUTXM,
        Wr{PK_F(4), 1, PK_W(1)}, Wr{PK_F(8), 1, PK_W(5)}, Wr{PK_F(5), 2, PK_Z(1)},
        Wr{PK_W(3), 1, PK_W(2)}, Wr{PK_W(7), 1, PK_W(6)}, Wr{PK_W(6), 2, PK_Z(2)}, Wr{PK_W(8), 2, PK_Z(4)},
        Wr{PK_F(1), 2, PK_Z(3)}}};
    case L_ESB: return WallProg{3, {PK_XP, PK_YM(8), PK_ZM(16)}, 12, {PK_OUTXP,
        Wr{PK_F(3), 1, PK_W(2)}, Wr{PK_F(7), 1, PK_W(6)}, Wr{PK_F(6), 2, PK_Z(2)},
        Wr{PK_W(4), 1, PK_W(1)}, Wr{PK_W(8), 1, PK_W(5)}, Wr{PK_W(5), 2, PK_Z(1)}, Wr{PK_W(7), 2, PK_Z(3)},
        Wr{PK_F(2), 2, PK_Z(4)}}};
    default: return WallProg{0, {}, 0, {}};
  }
}

// Bit masks (over the 24 value slots) the program reads before writing.
PIKS_HD constexpr uint32_t wall_prog_defined(const WallProg& P) {
  uint32_t m = 0;
  for (int a = 0; a < P.nface; ++a)
    for (int q = 0; q < 4; ++q) m |= 1u << (P.face[a].store + P.face[a].in[q]);
  return m;
}
PIKS_HD constexpr uint32_t wall_prog_written(const WallProg& P) {
  uint32_t m = wall_prog_defined(P);
  for (int t = 0; t < P.nwr; ++t) m |= 1u << P.wr[t].dst;
  return m;
}
PIKS_HD constexpr bool wall_prog_reads_only_defined(const WallProg& P) {
  uint32_t m = wall_prog_defined(P);
  for (int t = 0; t < P.nwr; ++t) {
    if (!((m >> P.wr[t].src) & 1u)) return false;
    m |= 1u << P.wr[t].dst;
  }
  return true;
}

constexpr int kWallLabels[26] = {L_E, L_W, L_N, L_S, L_F, L_B, L_EN, L_WN, L_ES, L_WS,
                                 L_NF, L_NB, L_SF, L_SB, L_EF, L_EB, L_WF, L_WB,
                                 L_ENF, L_WNF, L_WSF, L_ESF, L_ENB, L_WNB, L_WSB, L_ESB};

// Every program reads only values it produced itself, so a wall's own f and
// fw entries never need loading: they are only written.
constexpr bool all_programs_self_contained() {
  for (int t = 0; t < 26; ++t)
    if (!wall_prog_reads_only_defined(wall_prog(kWallLabels[t]))) return false;
  return true;
}
static_assert(all_programs_self_contained(), "a wall program reads a slot it did not define");

// A switch source slot is never itself a switch target, so the switches can be
// applied after the sweep in any order (checked over all labels and octants).
constexpr bool switch_sources_are_not_targets() {
  for (int o = 0; o < 8; ++o)
    for (int axis = 1; axis <= 2; ++axis)
      for (int t = 0; t < 26; ++t) {
        int lab = kWallLabels[t];
        int s = switch_source(o, axis, lab);
        if (s >= 1 && s <= 8) {
          int o2 = s - 1;
          if (switch_source(o2, 1, lab) != 0 || switch_source(o2, 2, lab) != 0) return false;
        }
      }
  return true;
}
static_assert(switch_sources_are_not_targets(), "switch source is also a switch target");

}  // namespace piks
