// One full red-black Gauss-Seidel iteration of a level in one launch: colour
// c1, then colour c2 = 1 - c1 (the two passes of GMGSolver.cu:216-269), on the
// colour-split V-cycle arrays (device code, included by kernels.hpp).
//
// The second pass of a tile needs, next to each of its faces, the values the
// first pass gave the face neighbour's cells of colour c1.  Those cells are
// updated here once more, with the arithmetic their own tile's wavefront uses
// (gs_quad / cell_row_m): the neighbour's face coefficients and neighbour
// table come from its record (Ctx::tiles), its face masks and right-hand side
// from memory, and the colour-c2 values its row reads (its face layer, its
// next layer, the tiles beside it across its own faces, and this tile's face
// layer) are all old values, as in the first pass.  A GHOST or missing
// neighbour is not smoothed: its values stay as stored (0 for ghosts in the
// down-sweep, 0 outside the grid).  Every value stored is therefore the value
// the two separate passes store, in every storage type (each first-pass
// value is rounded as its store would round it before the second pass reads
// it).  The kernel reads one array and writes another: every tile of the
// launch reads its neighbours' old values of both colours, so none may be
// overwritten during it.  Bytes per cell: x read and written once, the rhs and face masks read
// once (3s + t), against 2 (2s + t) for the two passes; the halo loads (about
// two face layers per face) mostly hit the L2 of the XCD that works the
// neighbours.
// Included inside namespace cg.
#pragma once

// The fields of a face neighbour's record the halo update reads.
struct PairNbr {
  float bc[7];        // its face coefficients (smoothing view at its level)
  int32_t nbr[6];     // its face neighbours (-1: none)
  uint32_t ntypes;    // their types, 4 bits each
};
struct PairLDS {
  float4 q[100];         // colour-c2 quads of the tile's pencils and the halo ring (pass 1), then colour-c1 (pass 2)
  float4 corner[4][2];   // ring group g, end e: the pencil beyond the ring's end (the neighbour's edge neighbour), colour c2
  float pl[2][10][10];   // z- plane (cells (i,j,7) of the z- neighbour) and z+ plane ((i,j,0) of the z+ one), index [i+1][j+1]
  PairNbr nb[6];
};

// ring group g (0: x-, 1: x+, 2: y-, 3: y+) -> (ring_face, kernels.hpp) the ring
// pencil's (i,j) in the neighbour, its deeper pencil, the tile's own adjacent
// pencil and the LDS index of ring position t
__device__ __forceinline__ void ring_ij(int g, int t, int& ri, int& rj) {
  ri = g == 0 ? 7 : (g == 1 ? 0 : t);
  rj = g == 2 ? 7 : (g == 3 ? 0 : t);
}
__device__ __forceinline__ int ring_deep(int g, int t) {
  return g == 0 ? (6 * 8 + t) : (g == 1 ? (1 * 8 + t) : (g == 2 ? (t * 8 + 6) : (t * 8 + 1)));
}
__device__ __forceinline__ int ring_slot(int g, int t) {
  return g == 0 ? (t + 1) : (g == 1 ? (91 + t) : (g == 2 ? ((t + 1) * 10) : ((t + 1) * 10 + 9)));
}
// the in-plane directions along ring group g: (face of the neighbour toward
// t = -1, toward t = 8)
__device__ __forceinline__ void ring_side_faces(int g, int& fm, int& fp) {
  fm = g < 2 ? 1 : 0;  // x rings run along y, y rings along x
  fp = g < 2 ? 4 : 3;
}

// Whether the values of tile n, reached through a neighbour relation of type
// ty, are read as stored (else they count as 0: outside the grid, and on
// ghosts in the down-sweep).
__device__ __forceinline__ bool pair_readable(int n, uint8_t ty, bool ghost_zero) {
  return n >= 0 && !(ghost_zero && ty == T_GHOST);
}

// XI holds the values of the level's smoothed (LEAF|NONLEAF) tiles, XG those
// of its GHOST tiles; the results of the listed tiles go to XO, which must be
// neither (every tile of the launch reads its neighbours' old values while
// the others write theirs).
template <class V, bool DOT>
__global__ __launch_bounds__(64) void k_gs_pair(Ctx c, const TileRec* __restrict__ list, int n, const V* __restrict__ XI,
                                                const V* __restrict__ XG, V* __restrict__ XO, int c1, int flags) {
  TILE_KERNEL_PROLOGUE
  __shared__ PairLDS L;
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const int c2 = c1 ^ 1;
  const bool ghost_zero = flags & F_GHOST_ZERO;
  const V* R = vc_rhs<V>(c);
  const int self = T.self;
  const float h = T.h;
  const V* zp = zero_page<V>();
  // ---- round 1: every load that needs only the tile's record, all issued
  // before the first wait.  Values that do not exist or count as 0 (outside
  // the grid, ghosts in the down-sweep, lanes past the ring) are read from the
  // zero page; values only an inactive neighbour's update would read come
  // from the zero page too and are not used.
  // the face neighbours' records (lanes 0..5 keep theirs)
  const int lf = lane < 6 ? lane : 5;
  const int nbv = pick4(lane & 3, uni(T.nbr[0]), uni(T.nbr[1]), uni(T.nbr[2]), uni(T.nbr[3]));
  const int nb = lane < 4 ? nbv : (lane == 4 ? uni(T.nbr[4]) : uni(T.nbr[5]));
  const bool nact = nb >= 0 && (tr_ntype(T, lf) & (T_LEAF | T_NONLEAF));
  const TileRec* nrec = c.tiles + (nact ? nb : self);
  PairNbr pn;
#pragma unroll
  for (int e = 0; e < 7; e++) pn.bc[e] = nrec->bc[e];
#pragma unroll
  for (int e = 0; e < 6; e++) pn.nbr[e] = nrec->nbr[e];
  pn.ntypes = nrec->ntypes;
  const int32_t* dp = (flags & F_CHECK_DONE) ? &c.ps[T.prob].done : reinterpret_cast<const int32_t*>(c.fmask);
  const int32_t done = *dp;
  const V* own = XI + size_t(self) * 512;
  typename RawV<V>::Q qxa = load_q_raw(own + (c1 << 8) + lane * 4);  // colour c1 (old)
  typename RawV<V>::Q qxb = load_q_raw(own + (c2 << 8) + lane * 4);  // colour c2 (old)
  typename RawV<V>::Q qba = load_q_raw(R + size_t(self) * 512 + (c1 << 8) + lane * 4);
  typename RawV<V>::Q qbb = load_q_raw(R + size_t(self) * 512 + (c2 << 8) + lane * 4);
  const uint64_t mask = load_ct8(c.fmask + size_t(self) * 512 + lane * 8);
  // the r pencil of a LEAF tile, for the r.z partial (DOT)
  float4 r0 = make_float4(0.f, 0.f, 0.f, 0.f), r1v = r0;
  if (DOT) {
    const float* rr = tr_type(T) == T_LEAF ? c.r + size_t(self) * 512 : zero_page<float>();
    const float4* rpn = reinterpret_cast<const float4*>(rr + lane * 8);
    r0 = rpn[0];
    r1v = rpn[1];
  }
  // halo ring (lanes 0..31): the neighbour's face pencil, both colours, and,
  // where that neighbour is smoothed, what its colour-c1 update reads
  const int g = (lane >> 3) & 3, t = lane & 7;
  const int rf = ring_face(g);
  int ri, rj;
  ring_ij(g, t, ri, rj);
  const int rpl = ri * 8 + rj;
  const int rn = pick4(g, uni(T.nbr[0]), uni(T.nbr[3]), uni(T.nbr[1]), uni(T.nbr[4]));
  const uint8_t rty = tr_ntype(T, rf);
  const bool rok = lane < 32 && pair_readable(rn, rty, ghost_zero);
  const bool ract = lane < 32 && rn >= 0 && (rty & (T_LEAF | T_NONLEAF));  // implies rok
  const V* RT = rok ? (rty == T_GHOST ? XG : XI) + size_t(rn) * 512 : zp;
  const V* RB = rok ? R + size_t(rn) * 512 : zp;
  typename RawV<V>::Q qrc1 = load_q_raw(rok ? RT + (c1 << 8) + rpl * 4 : zp);
  typename RawV<V>::Q qrc2 = load_q_raw(rok ? RT + (c2 << 8) + rpl * 4 : zp);
  typename RawV<V>::Q qrb = load_q_raw(rok ? RB + (c1 << 8) + rpl * 4 : zp);
  const uint64_t rm = load_ct8(rok ? c.fmask + size_t(rn) * 512 + rpl * 8 : reinterpret_cast<const uint8_t*>(zp));
  const int dpl = ring_deep(g, t);
  typename RawV<V>::Q qrdeep = load_q_raw(rok ? RT + (c2 << 8) + dpl * 4 : zp);

  // z planes (every lane): cell (i,j,7) of the z- neighbour and (i,j,0) of
  // the z+ one (old values); the lane whose own cell next to a plane has
  // colour c2 also updates that plane cell (colour c1) and reads its row
  const int nzm = T.nbr[2], nzp = T.nbr[5];
  const uint8_t tzm = tr_ntype(T, 2), tzp = tr_ntype(T, 5);
  const bool okm = pair_readable(nzm, tzm, ghost_zero), okp = pair_readable(nzp, tzp, ghost_zero);
  const bool actm = nzm >= 0 && (tzm & (T_LEAF | T_NONLEAF)), actp = nzp >= 0 && (tzp & (T_LEAF | T_NONLEAF));
  const bool upd_m = ((i + j) & 1) == c2;  // own (i,j,0) is colour c2: the z- plane cell (i,j,7) is colour c1
  const bool upd_p = !upd_m;              // own (i,j,7) is colour c2: the z+ plane cell (i,j,0) is colour c1
  const V* XM = okm ? (tzm == T_GHOST ? XG : XI) + size_t(nzm) * 512 : zp;
  const V* XP = okp ? (tzp == T_GHOST ? XG : XI) + size_t(nzp) * 512 : zp;
  typename RawV<V>::S svm = load_s_raw(okm ? XM + split_cell(i, j, 7) : zp);
  typename RawV<V>::S svp = load_s_raw(okp ? XP + split_cell(i, j, 0) : zp);
  // the updating lane's row data: deeper cell, rhs, face mask (used when that
  // plane's tile is smoothed, which implies it is readable)
  const bool oku = upd_m ? okm : okp;
  const int tu = upd_m ? nzm : nzp;
  const int ku = upd_m ? 7 : 0, kd = upd_m ? 6 : 1;
  typename RawV<V>::S szdeep = load_s_raw(oku ? (upd_m ? XM : XP) + split_cell(i, j, kd) : zp);
  typename RawV<V>::S szb = load_s_raw(oku ? R + size_t(tu) * 512 + split_cell(i, j, ku) : zp);
  const uint32_t zmask = *(oku ? c.fmask + size_t(tu) * 512 + lane * 8 + ku : reinterpret_cast<const uint8_t*>(zp));

  // ---- one wait, then round 1's values into LDS
  hold_all(qxa);
  hold_all(qxb);
  hold_all(qba);
  hold_all(qbb);
  hold_all(qrc1);
  hold_all(qrc2);
  hold_all(qrb);
  hold_all(qrdeep);
  hold_all(svm);
  hold_all(svp);
  hold_all(szdeep);
  hold_all(szb);
  const float4 xa = quad_f<V>(qxa), xb = quad_f<V>(qxb), ba = quad_f<V>(qba), bb = quad_f<V>(qbb);
  const float4 rc1 = quad_f<V>(qrc1), rc2 = quad_f<V>(qrc2), rb = quad_f<V>(qrb), rdeep = quad_f<V>(qrdeep);
  const float vm = single_f<V>(svm), vp = single_f<V>(svp);
  const float zdeep = single_f<V>(szdeep), zb = single_f<V>(szb);
  if (lane < 6) L.nb[lane] = pn;
  L.q[(i + 1) * 10 + (j + 1)] = xb;
  if (lane < 32) L.q[ring_slot(g, t)] = rc2;
  L.pl[0][i + 1][j + 1] = vm;
  L.pl[1][i + 1][j + 1] = vp;
  wave_sync();

  // ---- round 2 (needs the neighbours' records): the cells beyond the ends of
  // the ring's face layer and beyond the planes' edges, and the ring's k ends
  {
    // ring ends: lanes t == 0 and t == 7 of each group, the pencil beside the
    // ring pencil across the neighbour's in-plane face
    int fm, fp;
    ring_side_faces(g, fm, fp);
    const PairNbr& rnb = L.nb[rf];
    const bool end = lane < 32 && ract && (t == 0 || t == 7);
    const int ef = t == 0 ? fm : fp;
    const int en = rnb.nbr[ef];
    const uint8_t ety = uint8_t((rnb.ntypes >> (4 * ef)) & 0xf);
    const bool eok = end && pair_readable(en, ety, ghost_zero);
    // the edge pencil: across y (x rings) it is (ri, 7 - ...) etc.
    const int ei = g < 2 ? ri : (t == 0 ? 7 : 0);
    const int ej = g < 2 ? (t == 0 ? 7 : 0) : rj;
    typename RawV<V>::Q qev =
        load_q_raw(eok ? (ety == T_GHOST ? XG : XI) + size_t(en) * 512 + (c2 << 8) + (ei * 8 + ej) * 4 : zp);
    // the ring pencil's k end: its colour-c1 cells at even k (p = 0) need
    // cell (ri,rj,-1) = (ri,rj,7) of its z- neighbour, else (ri,rj,8) = (ri,rj,0) of its z+ one
    const bool rp = ((c1 + ri + rj) & 1) != 0;
    const int kf = rp ? 5 : 2;
    const int kn = rnb.nbr[kf];
    const uint8_t kty = uint8_t((rnb.ntypes >> (4 * kf)) & 0xf);
    const bool kok = lane < 32 && ract && pair_readable(kn, kty, ghost_zero);
    typename RawV<V>::S skv =
        load_s_raw(kok ? (kty == T_GHOST ? XG : XI) + size_t(kn) * 512 + split_cell(ri, rj, rp ? 0 : 7) : zp);
    // plane edges: lanes 0..31 fetch one edge cell of each plane
    //   q = 0: (-1, t) from the plane tile's x- neighbour, cell (7,t,K); 1: (8, t) from x+, (0,t,K);
    //   2: (t, -1) from y-, (t,7,K); 3: (t, 8) from y+, (t,0,K)
    const int q = g;
    const int pf = q == 0 ? 0 : (q == 1 ? 3 : (q == 2 ? 1 : 4));
    const int pi = q == 0 ? 7 : (q == 1 ? 0 : t), pj = q == 2 ? 7 : (q == 3 ? 0 : t);
    const PairNbr& zmn = L.nb[2];
    const PairNbr& zpn = L.nb[5];
    const int emn = zmn.nbr[pf], epn = zpn.nbr[pf];
    const uint8_t emt = uint8_t((zmn.ntypes >> (4 * pf)) & 0xf), ept = uint8_t((zpn.ntypes >> (4 * pf)) & 0xf);
    const bool emok = lane < 32 && actm && pair_readable(emn, emt, ghost_zero);
    const bool epok = lane < 32 && actp && pair_readable(epn, ept, ghost_zero);
    typename RawV<V>::S semv =
        load_s_raw(emok ? (emt == T_GHOST ? XG : XI) + size_t(emn) * 512 + split_cell(pi, pj, 7) : zp);
    typename RawV<V>::S sepv =
        load_s_raw(epok ? (ept == T_GHOST ? XG : XI) + size_t(epn) * 512 + split_cell(pi, pj, 0) : zp);
    hold_all(qev);
    hold_all(skv);
    hold_all(semv);
    hold_all(sepv);
    const float4 ev = quad_f<V>(qev);
    const float kv = single_f<V>(skv), emv = single_f<V>(semv), epv = single_f<V>(sepv);
    if (end) L.corner[g][t == 0 ? 0 : 1] = ev;
    if (lane < 32) {
      const int a = q == 0 ? 0 : (q == 1 ? 9 : t + 1), b2 = q == 2 ? 0 : (q == 3 ? 9 : t + 1);
      L.pl[0][a][b2] = emv;
      L.pl[1][a][b2] = epv;
    }
    wave_sync();

    // ---- pass 1: the tile's colour-c1 cells
    const int pc = (i + 1) * 10 + (j + 1);
    const bool p1 = ((c1 + i + j) & 1) != 0;
    const float4 n1 = gs_quad(xa, xb, L.q[pc - 10], L.q[pc + 10], L.q[pc - 1], L.q[pc + 1], vm, vp, ba, quad_mask(mask, p1), T.bc,
                              h, i, j, p1);
    const float4 x1 = make_float4(vround<V>(n1.x), vround<V>(n1.y), vround<V>(n1.z), vround<V>(n1.w));

    // the ring pencil's colour-c1 cells as the neighbour's first pass leaves them
    float4 r1 = rc1;
    if (ract) {
      const int s = ring_slot(g, t);
      const int st = g < 2 ? 1 : 10;  // LDS step along the ring
      const float4 along_m = t == 0 ? L.corner[g][0] : L.q[s - st];
      const float4 along_p = t == 7 ? L.corner[g][1] : L.q[s + st];
      // the tile's own pencil next to the ring pencil (colour c2, old)
      const int os = g == 0 ? (10 + t + 1) : (g == 1 ? (80 + t + 1) : (g == 2 ? ((t + 1) * 10 + 1) : ((t + 1) * 10 + 8)));
      const float4 ownq = L.q[os];
      float4 XIM, XIP, XJM, XJP;
      if (g == 0) { XIM = rdeep; XIP = ownq; XJM = along_m; XJP = along_p; }
      else if (g == 1) { XIM = ownq; XIP = rdeep; XJM = along_m; XJP = along_p; }
      else if (g == 2) { XIM = along_m; XIP = along_p; XJM = rdeep; XJP = ownq; }
      else { XIM = along_m; XIP = along_p; XJM = ownq; XJP = rdeep; }
      const float4 nr = gs_quad(rc1, rc2, XIM, XIP, XJM, XJP, rp ? 0.f : kv, rp ? kv : 0.f, rb, quad_mask(rm, rp), rnb.bc, h, ri,
                                rj, rp);
      r1 = make_float4(vround<V>(nr.x), vround<V>(nr.y), vround<V>(nr.z), vround<V>(nr.w));
    }
    // the plane cell next to the lane's colour-c2 end cell, as its first pass leaves it
    float kh;  // value of (i,j,-1) (upd_m) or (i,j,8) (upd_p) for the second pass
    {
      const int pw = upd_m ? 0 : 1;
      const bool act = upd_m ? actm : actp;
      const float x0 = upd_m ? vm : vp;
      kh = x0;
      if (act) {
        const PairNbr& zn = L.nb[upd_m ? 2 : 5];
        const float own_end = upd_m ? xb.x : xb.w;  // own (i,j,0) / (i,j,7), the lane's colour-c2 end cell
        const float bi = zn.bc[6];
        const float xn[6] = {L.pl[pw][i][j + 1], L.pl[pw][i + 2][j + 1], L.pl[pw][i + 1][j], L.pl[pw][i + 1][j + 2],
                             upd_m ? zdeep : own_end, upd_m ? own_end : zdeep};
        const float bc[6] = {i == 0 ? zn.bc[0] : bi, i == 7 ? zn.bc[3] : bi, j == 0 ? zn.bc[1] : bi,
                             j == 7 ? zn.bc[4] : bi, upd_m ? bi : zn.bc[2], upd_m ? zn.bc[5] : bi};
        float ax, dg;
        cell_row_m<true>(x0, zmask, xn, bc, h, ax, dg);
        float delta = 0.f;
        if ((zmask & M_INT) && dg != 0.f) delta = (zb - ax) / dg;
        kh = vround<V>(x0 + delta);
      }
    }
    wave_sync();  // every read of the colour-c2 quads is done
    L.q[pc] = x1;
    if (lane < 32) L.q[ring_slot(g, t)] = r1;
    wave_sync();
    if ((flags & F_CHECK_DONE) && done) return;

    // ---- pass 2: the tile's colour-c2 cells
    const bool p2 = !p1;
    const float4 n2 = gs_quad(xb, x1, L.q[pc - 10], L.q[pc + 10], L.q[pc - 1], L.q[pc + 1], upd_m ? kh : 0.f,
                              upd_p ? kh : 0.f, bb, quad_mask(mask, p2), T.bc, h, i, j, p2);
    V* tile = XO + size_t(self) * 512;
    store4(tile + (c1 << 8) + lane * 4, x1);
    store4(tile + (c2 << 8) + lane * 4, n2);
    if (DOT && tr_type(T) == T_LEAF) {
      // z final: colour c2 as just stored, colour c1 from pass 1; r.z over the
      // lane's pencil in cell order, as k_cg_dot_rz sums it (r loaded in round 1)
      const float zc[4] = {vround<V>(n2.x), vround<V>(n2.y), vround<V>(n2.z), vround<V>(n2.w)};
      const float zo[4] = {x1.x, x1.y, x1.z, x1.w};
      const float rv[8] = {r0.x, r0.y, r0.z, r0.w, r1v.x, r1v.y, r1v.z, r1v.w};
      double a = 0.0;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const float z = ((k & 1) == int(p2)) ? zc[k >> 1] : zo[k >> 1];
        a += double(rv[k]) * double(z);
      }
      a = wave_sum(a);
      if (lane == 0) c.part[T.leaf_slot] = a;
    }
  }
}
