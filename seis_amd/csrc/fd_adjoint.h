// Exact transpose of one step of the Born scattered-field recursion, in
// gather form.
//
// The forward step (fd_cell.h) maps (cur, prev, injection) to the new level
// through: stencil (or ring formula) -> raw; one-way band update ow(raw, cur,
// prev); blend w*raw + (1-w)*ow.  Its transpose, run backwards in time, maps
// the adjoint levels lam^m (complete) and lam^{m+1} to lam^{m-1}:
//
//   lam^{m-1} = sum over cells c that read cur(.) at step m of
//               (adjoint of raw/ow of c at level m) * d(c)/d cur
//             + the same for prev(.) with level m+1 + record injection.
//
// Every term is a gather, so each cell of lam^{m-1} is a pure function of
// lam^m, lam^{m+1} and per-band-cell products of those levels, each a fixed
// coefficient times lam(b) (BandAdj below).  The scatter-form statement-by-statement transpose
// in oracle/seis_oracle.cpp (born_adjoint) is the checker.
//
// The reference has no transpose of its Born operator (its `adjoint`,
// fd.cpp:592-719, is an imaging operator); this is the operator pair that
// least-squares RTM needs.
#pragma once
#include "fd_cell.h"

namespace seis {

// Position of a padded cell relative to the absorbing band, with the same
// partition as processBoundary (fd.cpp:1092-1178) and final_value.
struct BandCell {
  int kind;    // 0 interior, 1 z side (top/bottom), 2 x side (left/right), 3 corner
  int s;       // sides: +1 when the inner neighbour is at +1 along the normal
  int sx, sz;  // corners: direction of the x and z neighbours that are read
};

SEIS_HD inline BandCell classify(int x, int z, int NX, int NZ, int pml) {
  const int dl = x, dr = NX - 1 - x, dtp = z, db = NZ - 1 - z;
  const int mn = imin(imin(dl, dr), imin(dtp, db));
  BandCell b{0, 0, 0, 0};
  if (mn >= pml) return b;
  const bool zs = (dtp == mn || db == mn), xs = (dl == mn || dr == mn);
  if (zs && xs) {
    b.kind = 3;
    b.sz = (dtp == mn) ? 1 : -1;
    b.sx = (dl == mn) ? 1 : -1;
  } else if (zs) {
    b.kind = 1;
    b.s = (dtp == mn) ? 1 : -1;
  } else {
    b.kind = 2;
    b.s = (dl == mn) ? 1 : -1;
  }
  return b;
}

// Products of every band cell b of an adjoint level lam, with
// at = -(1-w) lam(b) / den the adjoint of the one-way quantity t of b:
//   sides  : in = (-a1+a2+2a3) at   (raw(b_in) and prev(b) coefficient)
//            t  = -a3 at            (raw and prev tangent neighbours)
//            c  = -2 a2 at          (cur(b) and cur(b_in))
//   corners: in = -b1 at (raw z neighbour), t = -b2 at (raw x neighbour),
//            c = -b3 at (cur(b))
template <typename T>
struct BandAdj {
  T in, t, c;
};

template <typename T>
SEIS_HD inline BandAdj<T> band_adj(const Consts<T>& k, const Model<T>& m, int x, int z, const BandCell& b, T lam) {
  BandAdj<T> r{T(0), T(0), T(0)};
  if (b.kind == 0) return r;
  const T v = m.vel(x, z);
  const T w = sponge_weight<T>(x, z, m.nx, m.nz, k.pml);
  if (b.kind == 3) {
    const T b3 = (b.sx == 1 && b.sz == 1) ? (T)(std::sqrt(2.0) / (double)v / (double)k.dt) : k.sq2 / v / k.dt;
    const T at = -(T(1) - w) * lam / (k.b1 + k.b2 + b3);
    r.in = -k.b1 * at;
    r.t = -k.b2 * at;
    r.c = -b3 * at;
    return r;
  }
  const T a1 = (b.kind == 1) ? k.a1z : k.a1x;
  const T a2 = k.a2b / v;
  const T h = (b.kind == 1) ? k.dx : k.dz;
  const T a3 = v / T(4) / h / h;
  const T at = -(T(1) - w) * lam / (a1 + a2);
  r.in = (-a1 + a2 + T(2) * a3) * at;
  r.t = -a3 * at;
  r.c = -T(2) * a2 * at;
  return r;
}

// Which band-adjoint terms reach cell (x, z): one bit per (product, neighbour)
// that the raw adjoint and the one-way cur / prev terms add.  The set depends only
// on the cell's position, so a plan can compute it once per cell.
enum BandTerm : unsigned {
  // raw adjoint: band cells that read raw(x, z)
  BT_AIN_ZM = 1u << 0,     // ain(x, z-1): top side or top corner above
  BT_AT_ZM_XM = 1u << 1,   // at(x-1, z-1): top side (tangent reader)
  BT_AT_ZM_XP = 1u << 2,   // at(x+1, z-1)
  BT_AIN_ZP = 1u << 3,     // ain(x, z+1): bottom side or bottom corner below
  BT_AT_ZP_XM = 1u << 4,   // at(x-1, z+1)
  BT_AT_ZP_XP = 1u << 5,   // at(x+1, z+1)
  BT_AIN_XM = 1u << 6,     // ain(x-1, z): left side
  BT_ATC_XM = 1u << 7,     // at(x-1, z): left corner (x neighbour read)
  BT_AT_XM_ZM = 1u << 8,   // at(x-1, z-1): left side (tangent reader)
  BT_AT_XM_ZP = 1u << 9,   // at(x-1, z+1)
  BT_AIN_XP = 1u << 10,    // ain(x+1, z): right side
  BT_ATC_XP = 1u << 11,    // at(x+1, z): right corner
  BT_AT_XP_ZM = 1u << 12,  // at(x+1, z-1)
  BT_AT_XP_ZP = 1u << 13,  // at(x+1, z+1)
  // one-way cur: cur(b) by band cell b, cur(b_in) by the side cell b
  BT_AC_SELF = 1u << 14,
  BT_AC_ZM = 1u << 15,
  BT_AC_ZP = 1u << 16,
  BT_AC_XM = 1u << 17,
  BT_AC_XP = 1u << 18,
  // one-way prev: prev(b) and prev(b_in) of sides, tangent prev reads
  BT_AIN_SELF = 1u << 19,
  BT_PIN_ZM = 1u << 20,
  BT_PIN_ZP = 1u << 21,
  BT_PIN_XM = 1u << 22,
  BT_PIN_XP = 1u << 23,
  BT_TAN_XM = 1u << 24,  // at(x-1, z): z side on the same row
  BT_TAN_ZM = 1u << 25,  // at(x, z-1): x side in the same column
  BT_TAN_XP = 1u << 26,
  BT_TAN_ZP = 1u << 27,
};

SEIS_HD inline unsigned band_terms(int x, int z, int NX, int NZ, int pml) {
  unsigned mk = 0;
  BandCell b;
  if (z - 1 >= 0) {  // readers above: top sides and top corners
    b = classify(x, z - 1, NX, NZ, pml);
    if ((b.kind == 1 && b.s == 1) || (b.kind == 3 && b.sz == 1)) mk |= BT_AIN_ZM;
    if (b.kind == 1 && b.s == 1) mk |= BT_AC_ZM | BT_PIN_ZM;
    if (x - 1 >= 0 && (b = classify(x - 1, z - 1, NX, NZ, pml), b.kind == 1 && b.s == 1)) mk |= BT_AT_ZM_XM;
    if (x + 1 < NX && (b = classify(x + 1, z - 1, NX, NZ, pml), b.kind == 1 && b.s == 1)) mk |= BT_AT_ZM_XP;
    if (classify(x, z - 1, NX, NZ, pml).kind == 2) mk |= BT_TAN_ZM;
  }
  if (z + 1 < NZ) {  // readers below: bottom sides and bottom corners
    b = classify(x, z + 1, NX, NZ, pml);
    if ((b.kind == 1 && b.s == -1) || (b.kind == 3 && b.sz == -1)) mk |= BT_AIN_ZP;
    if (b.kind == 1 && b.s == -1) mk |= BT_AC_ZP | BT_PIN_ZP;
    if (x - 1 >= 0 && (b = classify(x - 1, z + 1, NX, NZ, pml), b.kind == 1 && b.s == -1)) mk |= BT_AT_ZP_XM;
    if (x + 1 < NX && (b = classify(x + 1, z + 1, NX, NZ, pml), b.kind == 1 && b.s == -1)) mk |= BT_AT_ZP_XP;
    if (classify(x, z + 1, NX, NZ, pml).kind == 2) mk |= BT_TAN_ZP;
  }
  if (x - 1 >= 0) {  // readers to the left: left sides and left corners
    b = classify(x - 1, z, NX, NZ, pml);
    if (b.kind == 2 && b.s == 1) mk |= BT_AIN_XM | BT_AC_XM | BT_PIN_XM;
    if (b.kind == 3 && b.sx == 1) mk |= BT_ATC_XM;
    if (b.kind == 1) mk |= BT_TAN_XM;
    if (z - 1 >= 0 && (b = classify(x - 1, z - 1, NX, NZ, pml), b.kind == 2 && b.s == 1)) mk |= BT_AT_XM_ZM;
    if (z + 1 < NZ && (b = classify(x - 1, z + 1, NX, NZ, pml), b.kind == 2 && b.s == 1)) mk |= BT_AT_XM_ZP;
  }
  if (x + 1 < NX) {  // readers to the right
    b = classify(x + 1, z, NX, NZ, pml);
    if (b.kind == 2 && b.s == -1) mk |= BT_AIN_XP | BT_AC_XP | BT_PIN_XP;
    if (b.kind == 3 && b.sx == -1) mk |= BT_ATC_XP;
    if (b.kind == 1) mk |= BT_TAN_XP;
    if (z - 1 >= 0 && (b = classify(x + 1, z - 1, NX, NZ, pml), b.kind == 2 && b.s == -1)) mk |= BT_AT_XP_ZM;
    if (z + 1 < NZ && (b = classify(x + 1, z + 1, NX, NZ, pml), b.kind == 2 && b.s == -1)) mk |= BT_AT_XP_ZP;
  }
  const BandCell self = classify(x, z, NX, NZ, pml);
  if (self.kind != 0) mk |= BT_AC_SELF;
  if (self.kind == 1 || self.kind == 2) mk |= BT_AIN_SELF;
  return mk;
}

// Derivative of the ring's extrapolated term 2 d1 - d2 (fd.cpp:1072-1091),
// d1 = c1 (u0 + u2 - 2 u1), d2 = c2 (u1 + u3 - 2 u2), with respect to u_j
// (j = 0..3): 2 c1 (e0 + e2 - 2 e1) - c2 (e1 + e3 - 2 e2), e_i = [j == i].
// Straight-line on purpose: the gfx950 backend of ROCm 7.2 placed the
// negation of the lane-divergent `switch` form (default case -c2) under the
// wrong execution mask, so lanes with j == 3 received +c2.  Every product
// here is exact (factors 0, +-1, +-2), so the values equal the branchy form's.
template <typename T>
SEIS_HD inline T extrap_coef(int j, T c1, T c2) {
  const T e0 = T(j == 0), e1 = T(j == 1), e2 = T(j == 2), e3 = T(j == 3);
  return T(2) * c1 * (e0 + e2 - T(2) * e1) - c2 * (e1 + e3 - T(2) * e2);
}

SEIS_HD inline bool is_ring(int x, int z, int NX, int NZ) { return x == 0 || z == 0 || x == NX - 1 || z == NZ - 1; }

SEIS_HD inline bool high_order(int x, int z, int NX, int NZ) { return x >= 4 && x <= NX - 5 && z >= 4 && z <= NZ - 5; }

// lam^{m-1}(x, z) from level m and level m+1, without the record injection.
// ar1 / ar2 give the raw adjoint of a cell at level m / m+1 (w lam plus the
// products of the band cells that read its raw value), `ow` the cell's
// one-way terms: the adjoint of the one-way updates' reads of cur(x, z) at
// level m plus those of prev(x, z) at level m+1 (zero unless edge_dist <= pml).
// Also returns the adjoint of the raw value of the cell at level m (the
// gradient's imaging factor).
template <typename T, class AR1, class AR2>
SEIS_HD inline T adj_value(const Consts<T>& k, const Model<T>& m, int NX, const AR1& ar1, const AR2& ar2, int x,
                           int z, T* raw_adj_out, T ow) {
  const int NZ = k.NZ;
  auto czc = [&](int zz) { return imin(imax(zz, 0), NZ - 1); };
  auto cxc = [&](int xx) { return imin(imax(xx, 0), NX - 1); };
  auto vt2 = [&](T v) {
    const T a = k.dt * v;
    return a * a;
  };
  // Every read first, at positions clamped into the grid; the sums below add
  // the terms whose conditions hold, in a fixed order (no read waits on a
  // branch).  Neighbour j = 2 (d - 1) + (e > 0) is (x, z - e d) for the z
  // terms and (x - e d, z) for the x terms, d = 1..4, e = -1, 1.
  const T a_self = ar1(x, z), v_self = m.vel(x, z);
  T az[8], vz[8], ax[8], vx[8];
  for (int d = 1; d <= 4; d++) {
    for (int e = -1; e <= 1; e += 2) {
      const int j = 2 * (d - 1) + (e > 0);
      az[j] = ar1(x, czc(z - e * d));
      vz[j] = m.vel(x, czc(z - e * d));
      ax[j] = ar1(cxc(x - e * d), z);
      vx[j] = m.vel(cxc(x - e * d), z);
    }
  }
  const T top_a = ar1(x, 0), top_v1 = m.vel(x, 1), top_v2 = m.vel(x, 2);
  const T bot_a = ar1(x, NZ - 1), bot_v1 = m.vel(x, NZ - 2), bot_v2 = m.vel(x, NZ - 3);
  const T lft_a = ar1(0, z), lft_v1 = m.vel(1, z), lft_v2 = m.vel(2, z);
  const T rgt_a = ar1(NX - 1, z), rgt_v1 = m.vel(NX - 2, z), rgt_v2 = m.vel(NX - 3, z);
  const T a_prev = ar2(x, z);
  T acc = T(0);
  if (raw_adj_out) *raw_adj_out = a_self;
  // (i) stencil reads of cur by non-ring cells
  if (!is_ring(x, z, NX, NZ)) {
    const T v2 = vt2(v_self);
    const T az0 = a_self * (v2 * k.idz2), ax0 = a_self * (v2 * k.idx2);
    acc += T(2) * a_self;
    if (high_order(x, z, NX, NZ)) {
      acc += az0 * k.c[0] + ax0 * k.c[0];
    } else {
      acc -= T(2) * az0;
      acc -= T(2) * ax0;
    }
  }
  for (int d = 1; d <= 4; d++) {
    for (int e = -1; e <= 1; e += 2) {
      const int j = 2 * (d - 1) + (e > 0);
      // z neighbour c = (x, z - e d) reads cur(x, z) at offset e d
      const int cz = z - e * d;
      if (cz >= 1 && cz <= NZ - 2 && x >= 1 && x <= NX - 2) {
        const bool hi = high_order(x, cz, NX, NZ);
        if (hi || d == 1) {
          const T a = az[j] * (vt2(vz[j]) * k.idz2);
          acc += hi ? a * k.c[d] : a;
        }
      }
      const int cx = x - e * d;
      if (cx >= 1 && cx <= NX - 2 && z >= 1 && z <= NZ - 2) {
        const bool hi = high_order(cx, z, NX, NZ);
        if (hi || d == 1) {
          const T a = ax[j] * (vt2(vx[j]) * k.idx2);
          acc += hi ? a * k.c[d] : a;
        }
      }
    }
  }
  // (ii) ring formulas reading cur (fd.cpp:1060-1091)
  if (is_ring(x, z, NX, NZ)) {
    const bool xe = (x == 0 || x == NX - 1), ze = (z == 0 || z == NZ - 1);
    if (xe && ze) {
      acc += T(2) * a_self;
    } else {
      const T f = ze ? k.cx : (x == 0 ? k.cz : k.cx);
      acc += a_self * (T(2) - T(2) * ((v_self * v_self) * f));
    }
  }
  if (z == 0 || z == NZ - 1) {  // edge-parallel neighbours on the top/bottom rows: x-1 (j 1), x+1 (j 0)
    if (x - 1 >= 1 && x - 1 <= NX - 2) acc += ax[1] * ((vx[1] * vx[1]) * k.cx);
    if (x + 1 >= 1 && x + 1 <= NX - 2) acc += ax[0] * ((vx[0] * vx[0]) * k.cx);
  }
  if (x == 0 || x == NX - 1) {  // edge-parallel neighbours on the left/right columns: z-1 (j 1), z+1 (j 0)
    const T f = (x == 0) ? k.cz : k.cx;
    if (z - 1 >= 1 && z - 1 <= NZ - 2) acc += az[1] * ((vz[1] * vz[1]) * f);
    if (z + 1 >= 1 && z + 1 <= NZ - 2) acc += az[0] * ((vz[0] * vz[0]) * f);
  }
  // extrapolated normal differences: ring cell (x, 0) reads cur(x, 0..3), etc.
  // (the coefficient of cur at distance j from the edge; straight-line form)
  if (z <= 3) acc += top_a * extrap_coef<T>(z, (top_v1 * top_v1) * k.cz, (top_v2 * top_v2) * k.cz);
  if (NZ - 1 - z <= 3) acc += bot_a * extrap_coef<T>(NZ - 1 - z, (bot_v1 * bot_v1) * k.cz, (bot_v2 * bot_v2) * k.cz);
  if (x <= 3) acc += lft_a * extrap_coef<T>(x, (lft_v1 * lft_v1) * k.cx, (lft_v2 * lft_v2) * k.cx);
  if (NX - 1 - x <= 3) acc += rgt_a * extrap_coef<T>(NX - 1 - x, (rgt_v1 * rgt_v1) * k.cx, (rgt_v2 * rgt_v2) * k.cx);
  // (iii) -prev of the stencil and ring base formula at level m+1, (iv) the
  // one-way reads of cur and prev
  acc -= a_prev;
  acc += ow;
  return acc;
}

}  // namespace seis
