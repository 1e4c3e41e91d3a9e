// HIP kernels of the MGPCG path for gfx950 (device code, included by mgpcg.hip).
//
// Layout in HBM (structure of arrays, one problem or many in one pool):
//   channel c of tile t, cell (i,j,k):  ch[c][t*512 + (i<<6 | j<<3 | k)]  (FP32)
//   cell types:                        ctype[t*512 + ...]                (u8)
//   per-tile table DevTile (64 B), per-problem scalars ProbState.
// Execution: one 64-lane wavefront per 8^3 tile.  Lane l = (i,j) = (l>>3, l&7)
// owns the k-pencil of 8 cells at offsets l*8..l*8+7 (32 contiguous bytes), so a
// wavefront reads or writes a whole channel of a tile as 2 KB of coalesced
// 16-byte accesses.  Neighbours in k are in the lane's registers; neighbours in
// i and j come from an LDS copy of the pencils plus a one-pencil halo ring
// loaded from the face-neighbour tiles; the two k-halo cells are read straight
// into registers.
//
// Floating-point operations are written out in the reference's order with
// explicit fmaf where the reference's compiler fuses a multiply-add, and the
// library is compiled with -ffp-contract=off, so every kernel reproduces the
// CPU transcription bit for bit.  FP64 dot products use one fixed summation
// order (per lane over its 8 cells, a xor-butterfly across the wavefront, then
// per problem over the tile partials the same way).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cg {

enum : uint8_t { T_LEAF = 1, T_NONLEAF = 2, T_GHOST = 4 };
enum : uint8_t { C_DIR = 1, C_INT = 2, C_NEU = 4 };

struct alignas(16) DevTile {
  int32_t nbr[6];     // same-level face neighbours x-,y-,z-,x+,y+,z+ (-1: none)
  int32_t parent;     // -1 on level 0
  int32_t prob;       // problem index in the pool
  int32_t leaf_slot;  // LEAF: slot of its dot partial; -1 otherwise
  int32_t gslot;      // GHOST: slot in the 64-value ghost buffers; -1 otherwise
  int32_t gchild;     // LEAF with ghost children: row of the 8-entry ghost-child table; -1 otherwise
  float h;            // cell size of the tile's level
  uint8_t level, type, octant, pad0;
  uint8_t nbr_type[6];
  uint8_t pad1[6];
};
static_assert(sizeof(DevTile) == 64, "DevTile is 64 bytes");

struct alignas(16) ProbState {
  double gamma, gamma_old, alpha, beta;
  double bb, thr, rr, pAp;
  double mean, tol2;
  int32_t iter, done, converged, max_iters;
  uint32_t counter;
  int32_t leaf_begin, leaf_count, pure_neumann;
};

// Channel pointers and tables passed to every kernel by value.
struct Ctx {
  const DevTile* tiles;
  const int32_t* gchild;  // 8 ghost-child tile indices (or -1) per row
  const int32_t* gcslot;  // the same children's ghost slots (or -1)
  const uint8_t* ctype;
  float* x;      // pressure / CG x
  float* r;      // rhs b, then residual r
  float* p;      // CG direction
  float* ap;     // A p
  float* z;      // preconditioned residual (V-cycle x)
  float* rhsv;   // V-cycle rhs of NONLEAF tiles
  float* zb;     // damped-Jacobi ping-pong buffer
  float* gsum;   // 64 per ghost: 2x2x2 sums of the ghost rows of A p
  float* gext;   // 64 per ghost: restricted ghost residual (V-cycle)
  ProbState* ps;
  double* part;  // per-leaf-slot dot partials
  double* part2; // second partial stream (mean counts)
};

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__host__ __device__ __forceinline__ uint8_t subtree_type(uint8_t type, int level, int sl) {
  return (level == sl && type == T_NONLEAF) ? T_LEAF : type;
}
// Face coefficient (times h^2 is 1/h^2 on LEAF-LEAF faces, (2/3)/h^2 on
// LEAF-GHOST faces, 0 otherwise) before the NEUMANN cut: GMGSolver.cu:26-39.
__host__ __device__ __forceinline__ float base_coeff(uint8_t t0, uint8_t t1, float cll, float clg) {
  bool both = (t0 & T_LEAF) && (t1 & T_LEAF);
  bool lg = ((t0 & T_LEAF) && (t1 & T_GHOST)) || ((t0 & T_GHOST) && (t1 & T_LEAF));
  return both ? cll : (lg ? clg : 0.f);
}
__device__ __forceinline__ uint8_t ct_at(uint64_t w, int k) { return uint8_t(w >> (8 * k)); }

__device__ __forceinline__ void load8(const float* __restrict__ p, float v[8]) {
  float4 a = reinterpret_cast<const float4*>(p)[0];
  float4 b = reinterpret_cast<const float4*>(p)[1];
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void store8(float* __restrict__ p, const float v[8]) {
  reinterpret_cast<float4*>(p)[0] = make_float4(v[0], v[1], v[2], v[3]);
  reinterpret_cast<float4*>(p)[1] = make_float4(v[4], v[5], v[6], v[7]);
}
__device__ __forceinline__ uint64_t load_ct8(const uint8_t* __restrict__ p) {
  return *reinterpret_cast<const uint64_t*>(p);
}

// Wavefront sum of a double in the fixed xor-butterfly order (masks 32..1);
// every lane ends with the same bits.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = v + __shfl_xor(v, m, 64);
  return v;
}

// Sum of a problem's per-tile partials in the fixed order: lane q adds
// partials q, q+64, q+128, ... in sequence (loads issued 8 at a time), then the
// wavefront butterfly.
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int begin, int count) {
  double a = 0.0;
  int q = lane_id();
  for (; q + 7 * 64 < count; q += 8 * 64) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = part[begin + q + u * 64];
#pragma unroll
    for (int u = 0; u < 8; u++) a += v[u];
  }
  for (; q < count; q += 64) a += part[begin + q];
  return wave_sum(a);
}

// Block -> worklist position.  Consecutive workgroups go round-robin to the 8
// XCDs; giving each XCD a contiguous (Morton-ordered) run of tiles keeps the
// face halos of a tile in the L2 of the XCD that reads them.
__device__ __forceinline__ int work_index(int n, int xcd_swizzle) {
  int b = blockIdx.x;
  if (!xcd_swizzle) return b;
  int per = (int(gridDim.x) + 7) >> 3;
  return (b & 7) * per + (b >> 3);
}

// ---------------------------------------------------------------------------
// Tile + halo staging
// ---------------------------------------------------------------------------
struct TileLDS {
  float4 xs[2][100];   // pencils (i,j), i,j in -1..8, index (i+1)*10 + (j+1); [0]: k 0..3, [1]: k 4..7
  uint64_t cs[100];    // cell types of those pencils
};

struct Stencil {
  float own[8];
  uint64_t ct;         // own cell types
  float xkm, xkp;      // k-halo values (z- neighbour cell (i,j,7), z+ neighbour cell (i,j,0))
  uint8_t ckm, ckp;
  float b_int, b_f[6]; // base coefficients: inside the tile and across face f
  float h;
};

// Stage tile t's X values (zeros where `zero` says so) and cell types.
//   zero_all:   treat X as 0 everywhere (first sweep from a zero guess)
//   ghost_zero: treat X of GHOST tiles as 0 (downstroke of the V-cycle)
//   sl:         subtree level of the tile-type view (-1: plain types)
// X holds active tiles, XG ghost tiles (they differ only for the damped-Jacobi
// ping-pong, where ghost values stay in the primary buffer).
__device__ __forceinline__ void stage_tile(const Ctx& c, int t, const DevTile& T, const float* __restrict__ X,
                                           const float* __restrict__ XG, bool zero_all, bool ghost_zero, int sl,
                                           TileLDS& L, Stencil& S) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t base = size_t(t) * 512 + lane * 8;
  bool zero_own = zero_all || (ghost_zero && T.type == T_GHOST);
  if (zero_own) { for (int k = 0; k < 8; k++) S.own[k] = 0.f; }
  else load8((T.type == T_GHOST ? XG : X) + base, S.own);
  S.ct = load_ct8(c.ctype + base);
  const int pc = (i + 1) * 10 + (j + 1);
  L.xs[0][pc] = make_float4(S.own[0], S.own[1], S.own[2], S.own[3]);
  L.xs[1][pc] = make_float4(S.own[4], S.own[5], S.own[6], S.own[7]);
  L.cs[pc] = S.ct;

  // halo pencils: x- (i==0), x+ (i==7), y- (j==0), y+ (j==7)
  auto halo = [&](int f, int src_off, int dst) {
    int n = T.nbr[f];
    float v[8];
    uint64_t w;
    if (n < 0) {
      for (int k = 0; k < 8; k++) v[k] = 0.f;
      w = 0x0101010101010101ull;  // DIRICHLET outside the grid (GMGSolver.cu:120-122)
    } else {
      const size_t nb = size_t(n) * 512 + src_off;
      bool gh = T.nbr_type[f] == T_GHOST;
      bool z = zero_all || (ghost_zero && gh);
      if (z) { for (int k = 0; k < 8; k++) v[k] = 0.f; }
      else load8((gh ? XG : X) + nb, v);
      w = load_ct8(c.ctype + nb);
    }
    L.xs[0][dst] = make_float4(v[0], v[1], v[2], v[3]);
    L.xs[1][dst] = make_float4(v[4], v[5], v[6], v[7]);
    L.cs[dst] = w;
  };
  if (i == 0) halo(0, (7 * 8 + j) * 8, 0 * 10 + (j + 1));
  if (i == 7) halo(3, (0 * 8 + j) * 8, 9 * 10 + (j + 1));
  if (j == 0) halo(1, (i * 8 + 7) * 8, (i + 1) * 10 + 0);
  if (j == 7) halo(4, (i * 8 + 0) * 8, (i + 1) * 10 + 9);
  // k halos
  {
    int n = T.nbr[2];
    if (n < 0) { S.xkm = 0.f; S.ckm = C_DIR; }
    else {
      size_t o = size_t(n) * 512 + lane * 8 + 7;
      bool gh = T.nbr_type[2] == T_GHOST;
      bool z = zero_all || (ghost_zero && gh);
      S.xkm = z ? 0.f : (gh ? XG : X)[o];
      S.ckm = c.ctype[o];
    }
    n = T.nbr[5];
    if (n < 0) { S.xkp = 0.f; S.ckp = C_DIR; }
    else {
      size_t o = size_t(n) * 512 + lane * 8;
      bool gh = T.nbr_type[5] == T_GHOST;
      bool z = zero_all || (ghost_zero && gh);
      S.xkp = z ? 0.f : (gh ? XG : X)[o];
      S.ckp = c.ctype[o];
    }
  }
  // coefficients (GMGSolver.cu:26-39): 1/h^2 and (2/h/3)/h as the reference forms them
  S.h = T.h;
  float ooh = 1.0f / T.h;
  float cll = ooh * ooh;
  float clg = ((2.0f * ooh) / 3.0f) * ooh;
  uint8_t to = subtree_type(T.type, T.level, sl);
  S.b_int = base_coeff(to, to, cll, clg);
  for (int f = 0; f < 6; f++) {
    uint8_t tn = (T.nbr[f] < 0) ? to : subtree_type(T.nbr_type[f], T.level, sl);
    S.b_f[f] = base_coeff(to, tn, cll, clg);
  }
}

// A x (and its diagonal) on the lane's 8 cells after stage_tile + barrier.
// The six faces are visited in the reference's order x-, x+, y-, y+, z-, z+
// (GMGSolver.cu:62-73).
template <bool WANT_AX, bool WANT_D>
__device__ __forceinline__ void apply_pencil(const TileLDS& L, const Stencil& S, float ax[8], float dg[8]) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const int pc = (i + 1) * 10 + (j + 1);
  float xim[8], xip[8], xjm[8], xjp[8];
  {
    float4 a = L.xs[0][pc - 10], b = L.xs[1][pc - 10];
    xim[0] = a.x; xim[1] = a.y; xim[2] = a.z; xim[3] = a.w; xim[4] = b.x; xim[5] = b.y; xim[6] = b.z; xim[7] = b.w;
    a = L.xs[0][pc + 10]; b = L.xs[1][pc + 10];
    xip[0] = a.x; xip[1] = a.y; xip[2] = a.z; xip[3] = a.w; xip[4] = b.x; xip[5] = b.y; xip[6] = b.z; xip[7] = b.w;
    a = L.xs[0][pc - 1]; b = L.xs[1][pc - 1];
    xjm[0] = a.x; xjm[1] = a.y; xjm[2] = a.z; xjm[3] = a.w; xjm[4] = b.x; xjm[5] = b.y; xjm[6] = b.z; xjm[7] = b.w;
    a = L.xs[0][pc + 1]; b = L.xs[1][pc + 1];
    xjp[0] = a.x; xjp[1] = a.y; xjp[2] = a.z; xjp[3] = a.w; xjp[4] = b.x; xjp[5] = b.y; xjp[6] = b.z; xjp[7] = b.w;
  }
  const uint64_t cim = L.cs[pc - 10], cip = L.cs[pc + 10], cjm = L.cs[pc - 1], cjp = L.cs[pc + 1];
  const float bim = (i == 0) ? S.b_f[0] : S.b_int;
  const float bip = (i == 7) ? S.b_f[3] : S.b_int;
  const float bjm = (j == 0) ? S.b_f[1] : S.b_int;
  const float bjp = (j == 7) ? S.b_f[4] : S.b_int;
  const float h = S.h;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint8_t c0 = ct_at(S.ct, k);
    const float x0 = S.own[k];
    const uint8_t ckm = (k == 0) ? S.ckm : ct_at(S.ct, k - 1);
    const uint8_t ckp = (k == 7) ? S.ckp : ct_at(S.ct, k + 1);
    const float xkm = (k == 0) ? S.xkm : S.own[k - 1];
    const float xkp = (k == 7) ? S.xkp : S.own[k + 1];
    const float bkm = (k == 0) ? S.b_f[2] : S.b_int;
    const float bkp = (k == 7) ? S.b_f[5] : S.b_int;
    const bool n0 = c0 & C_NEU;
    const float c1 = (n0 || (ct_at(cim, k) & C_NEU)) ? 0.f : bim;
    const float c2 = (n0 || (ct_at(cip, k) & C_NEU)) ? 0.f : bip;
    const float c3 = (n0 || (ct_at(cjm, k) & C_NEU)) ? 0.f : bjm;
    const float c4 = (n0 || (ct_at(cjp, k) & C_NEU)) ? 0.f : bjp;
    const float c5 = (n0 || (ckm & C_NEU)) ? 0.f : bkm;
    const float c6 = (n0 || (ckp & C_NEU)) ? 0.f : bkp;
    const bool in = c0 & C_INT;
    if (WANT_AX) {
      float s = 0.f;
      s = __builtin_fmaf(c1, x0 - xim[k], s);
      s = __builtin_fmaf(c2, x0 - xip[k], s);
      s = __builtin_fmaf(c3, x0 - xjm[k], s);
      s = __builtin_fmaf(c4, x0 - xjp[k], s);
      s = __builtin_fmaf(c5, x0 - xkm, s);
      s = __builtin_fmaf(c6, x0 - xkp, s);
      ax[k] = in ? ((s * h) * h) * h : 0.f;
    }
    if (WANT_D) {
      float d = 0.f;
      d = d + c1; d = d + c2; d = d + c3; d = d + c4; d = d + c5; d = d + c6;
      dg[k] = in ? ((d * h) * h) * h : 0.f;
    }
  }
}

// rhs of the V-cycle on a LEAF|NONLEAF tile: NONLEAF tiles hold the restricted
// residual; LEAF tiles use r plus, in octants that have a ghost child, that
// ghost's restricted residual (the reference's rhs += ghost sum,
// GMGSolver.cu:330-331, on top of rhs = r from the V-cycle init, :393-400).
__device__ __forceinline__ void vcycle_rhs(const Ctx& c, int t, const DevTile& T, const float* __restrict__ f, float b[8]) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t base = size_t(t) * 512 + lane * 8;
  if (T.type == T_NONLEAF) { load8(c.rhsv + base, b); return; }
  load8(f + base, b);
  if (T.gchild < 0) return;
  for (int h = 0; h < 2; h++) {
    int o = (i >> 2) | ((j >> 2) << 1) | (h << 2);
    int gs = c.gcslot[T.gchild * 8 + o];
    if (gs < 0) continue;
    const float4 e = *reinterpret_cast<const float4*>(c.gext + size_t(gs) * 64 + (i & 3) * 16 + (j & 3) * 4);
    b[4 * h + 0] = b[4 * h + 0] + e.x;
    b[4 * h + 1] = b[4 * h + 1] + e.y;
    b[4 * h + 2] = b[4 * h + 2] + e.z;
    b[4 * h + 3] = b[4 * h + 3] + e.w;
  }
}

// ---------------------------------------------------------------------------
// Kernel flags
// ---------------------------------------------------------------------------
enum : int {
  F_ZERO_X = 1,      // X is implicitly zero (first sweep)
  F_GHOST_ZERO = 2,  // ghost X is zero (V-cycle downstroke)
  F_VCYCLE_RHS = 4,  // rhs via vcycle_rhs (else the rhs channel)
  F_VCYCLE_OUT = 8,  // restriction writes rhsv / gext (else the coarse channel)
  F_SWIZZLE = 16,    // XCD-aware tile order
  F_CHECK_DONE = 32, // skip tiles of converged problems
};

__device__ __forceinline__ bool prob_done(const Ctx& c, const DevTile& T, int flags) {
  return (flags & F_CHECK_DONE) && c.ps[T.prob].done;
}

// Each tile is worked by exactly one wavefront with its own LDS region, so
// staging only needs ordering inside the wavefront.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Per-tile bodies (one wavefront each), shared by the one-tile-per-workgroup
// kernels and the coarse-level kernel.
// ---------------------------------------------------------------------------
// Red-black Gauss-Seidel colour sweep of a LEAF|NONLEAF tile (GMGSolver.cu:216-269).
// An INTERIOR cell with D == 0 is left unchanged.
__device__ __forceinline__ void tile_gs(const Ctx& c, int t, const DevTile& T, int level, float* X, const float* F,
                                        int color, int flags, TileLDS& L) {
  Stencil S;
  stage_tile(c, t, T, X, X, flags & F_ZERO_X, flags & F_GHOST_ZERO, level, L, S);
  float b[8];
  if (flags & F_VCYCLE_RHS) vcycle_rhs(c, t, T, F, b);
  else load8(F + size_t(t) * 512 + lane_id() * 8, b);
  if (prob_done(c, T, flags)) return;
  wave_sync();
  float ax[8], dg[8];
  apply_pencil<true, true>(L, S, ax, dg);
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  float out[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    out[k] = S.own[k];
    if (((i + j + k) & 1) == color) {
      float delta = 0.f;
      if ((ct_at(S.ct, k) & C_INT) && dg[k] != 0.f) delta = (b[k] - ax[k]) / dg[k];
      out[k] = S.own[k] + delta;
    }
  }
  store8(X + size_t(t) * 512 + lane * 8, out);
}

// Damped-Jacobi sweep (the update of GMGSolver.cu:508-525 with the level's
// operator): Y = X + omega (b - A X) / D on INTERIOR cells, 0 where D == 0.
__device__ __forceinline__ void tile_jacobi(const Ctx& c, int t, const DevTile& T, int level, const float* X,
                                            const float* XG, float* Y, const float* F, float omega, int flags, TileLDS& L) {
  Stencil S;
  stage_tile(c, t, T, X, XG, flags & F_ZERO_X, flags & F_GHOST_ZERO, level, L, S);
  float b[8];
  if (flags & F_VCYCLE_RHS) vcycle_rhs(c, t, T, F, b);
  else load8(F + size_t(t) * 512 + lane_id() * 8, b);
  if (prob_done(c, T, flags)) return;
  wave_sync();
  float ax[8], dg[8];
  apply_pencil<true, true>(L, S, ax, dg);
  float out[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    out[k] = S.own[k];
    if (ct_at(S.ct, k) & C_INT) {
      float res = b[k] - ax[k];
      out[k] = (dg[k] == 0.f) ? 0.f : S.own[k] + (omega * res) / dg[k];
    }
  }
  store8(Y + size_t(t) * 512 + lane_id() * 8, out);
}

// 2x2x2 sums of a tile's pencil values, in the reference's (ii, jj, kk) order
// (GMGSolver.cu:319-328): lane q = (a,b,c) = (q>>4, (q>>2)&3, q&3) returns the
// sum of fine cells (2a+ii, 2b+jj, 2c+kk).  The values must be in LDS first.
__device__ __forceinline__ float octet_sum(const float* Ls, int q) {
  int a = q >> 4, bq = (q >> 2) & 3, cq = q & 3;
  float s = 0.f;
  for (int ii = 0; ii < 2; ii++)
    for (int jj = 0; jj < 2; jj++)
      for (int kk = 0; kk < 2; kk++) s += Ls[((2 * a + ii) * 8 + (2 * bq + jj)) * 8 + 2 * cq + kk];
  return s;
}
__device__ __forceinline__ size_t parent_cell(const DevTile& T, int q) {
  int a = q >> 4, bq = (q >> 2) & 3, cq = q & 3;
  int ox = T.octant & 1, oy = (T.octant >> 1) & 1, oz = (T.octant >> 2) & 1;
  return size_t(T.parent) * 512 + ((ox * 4 + a) << 6) + ((oy * 4 + bq) << 3) + (oz * 4 + cq);
}

// Residual and restriction of one tile of `level` to level-1
// (GMGSolver.cu:284-339).  V-cycle output: NONLEAF/LEAF fine tiles write the
// parent's rhsv (= sum * one_over_alpha); GHOST fine tiles write their 64 sums
// to gext (added to the parent LEAF's r when the parent level is smoothed).
// Channel output: parent channel = sum * one_over_alpha, or += sum for ghosts.
__device__ __forceinline__ void tile_restrict(const Ctx& c, int t, const DevTile& T, int level, const float* X,
                                              const float* F, float* CO, float one_over_alpha, int flags, TileLDS& L,
                                              float* res) {
  Stencil S;
  stage_tile(c, t, T, X, X, flags & F_ZERO_X, flags & F_GHOST_ZERO, level, L, S);
  const int lane = lane_id();
  float b[8];
  if (flags & F_VCYCLE_RHS) {
    if (T.type == T_GHOST) { for (int k = 0; k < 8; k++) b[k] = 0.f; }
    else vcycle_rhs(c, t, T, F, b);
  } else load8(F + size_t(t) * 512 + lane * 8, b);
  if (prob_done(c, T, flags)) return;
  wave_sync();
  float ax[8], dg[8];
  apply_pencil<true, false>(L, S, ax, dg);
  for (int k = 0; k < 8; k++) res[lane * 8 + k] = b[k] - ax[k];
  wave_sync();
  float s = octet_sum(res, lane);
  const bool ghost = (subtree_type(T.type, T.level, level) & T_GHOST);
  if (flags & F_VCYCLE_OUT) {
    if (ghost) c.gext[size_t(T.gslot) * 64 + lane] = s;
    else c.rhsv[parent_cell(T, lane)] = s * one_over_alpha;
  } else {
    size_t pc = parent_cell(T, lane);
    if (ghost) CO[pc] = CO[pc] + s;
    else CO[pc] = s * one_over_alpha;
  }
}

// Prolongation and update of one tile of `level` (GMGSolver.cu:345-376):
// ghosts take the parent value, others add prolong_coeff * parent value, on
// INTERIOR cells; other cells become 0.  Ghost results go to XG (the primary
// buffer) and active tiles to XO (a Jacobi ping-pong target).
__device__ __forceinline__ void tile_prolong(const Ctx& c, int t, const DevTile& T, int level, const float* XC,
                                             const float* XF, float* XO, float* XG, float pcoef, int flags) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t base = size_t(t) * 512 + lane * 8;
  int ox = T.octant & 1, oy = (T.octant >> 1) & 1, oz = (T.octant >> 2) & 1;
  const float4 pv = *reinterpret_cast<const float4*>(XC + size_t(T.parent) * 512 + ((ox * 4 + (i >> 1)) << 6) +
                                                     ((oy * 4 + (j >> 1)) << 3) + oz * 4);
  const float cz[4] = {pv.x, pv.y, pv.z, pv.w};
  const uint64_t ct = load_ct8(c.ctype + base);
  const bool ghost = subtree_type(T.type, T.level, level) & T_GHOST;
  float fx[8];
  if (ghost || (flags & F_ZERO_X)) { for (int k = 0; k < 8; k++) fx[k] = 0.f; }
  else load8(XF + base, fx);
  if (prob_done(c, T, flags)) return;
  float out[8];
  if (ghost) {
    for (int k = 0; k < 8; k++) out[k] = (ct_at(ct, k) & C_INT) ? cz[k >> 1] : 0.f;
    store8(XG + base, out);
  } else {
    for (int k = 0; k < 8; k++) out[k] = (ct_at(ct, k) & C_INT) ? __builtin_fmaf(pcoef, cz[k >> 1], fx[k]) : 0.f;
    store8(XO + base, out);
  }
}

// ---------------------------------------------------------------------------
// One-tile-per-workgroup kernels over a worklist
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_gs_color(Ctx c, const int32_t* __restrict__ list, int n, int level,
                                                 float* X, const float* F, int color, int flags) {
  int w = work_index(n, flags & F_SWIZZLE);
  if (w >= n) return;
  const int t = list[w];
  __shared__ TileLDS L;
  tile_gs(c, t, c.tiles[t], level, X, F, color, flags, L);
}
__global__ __launch_bounds__(64) void k_jacobi(Ctx c, const int32_t* __restrict__ list, int n, int level, const float* X,
                                               const float* XG, float* Y, const float* F, float omega, int flags) {
  int w = work_index(n, flags & F_SWIZZLE);
  if (w >= n) return;
  const int t = list[w];
  __shared__ TileLDS L;
  tile_jacobi(c, t, c.tiles[t], level, X, XG, Y, F, omega, flags, L);
}
__global__ __launch_bounds__(64) void k_restrict(Ctx c, const int32_t* __restrict__ list, int n, int level, const float* X,
                                                 const float* F, float* CO, float one_over_alpha, int flags) {
  int w = work_index(n, flags & F_SWIZZLE);
  if (w >= n) return;
  const int t = list[w];
  __shared__ TileLDS L;
  __shared__ float res[512];
  tile_restrict(c, t, c.tiles[t], level, X, F, CO, one_over_alpha, flags, L, res);
}
__global__ __launch_bounds__(64) void k_prolong(Ctx c, const int32_t* __restrict__ list, int n, int level,
                                                const float* XC, const float* XF, float* XO, float* XG, float pcoef, int flags) {
  int w = work_index(n, flags & F_SWIZZLE);
  if (w >= n) return;
  const int t = list[w];
  tile_prolong(c, t, c.tiles[t], level, XC, XF, XO, XG, pcoef, flags);
}

// ---------------------------------------------------------------------------
// Coarse levels 0..C of the V-cycle (smoothing, restriction, the bottom solve
// on level 0, prolongation) for one problem in one workgroup, with every tile
// of those levels (active and ghost, at most CL_MAX_TILES) resident in LDS and
// one thread per cell.  The host picks C so that these levels are tiny; one
// launch then replaces the ~10 launches per level a grid-wide phase each would
// need.  The per-cell arithmetic is the same as the tile kernels' (faces in
// the order x-, x+, y-, y+, z-, z+, fmaf, ((s*h)*h)*h, (b-Ax)/D), so results
// are identical; the bottom is GMGSolver.cu:432-433.
// ---------------------------------------------------------------------------
constexpr int CL_MAX_TILES = 16;
constexpr int CL_MAX_LEVELS = 8;
constexpr int CL_THREADS = 1024;
struct alignas(16) CoarseSlot {
  int32_t tile;      // tile index
  int32_t nbr[6];    // slot of the same-level face neighbour (x-,y-,z-,x+,y+,z+), -1: none
  int32_t parent;    // slot of the parent, -1 on level 0
  int32_t level;
  float h;
  uint8_t type, octant, pad[2];
  int32_t pad2;
};
static_assert(sizeof(CoarseSlot) == 48, "CoarseSlot is 48 bytes");
struct CoarsePlan {
  int C;                              // deepest level held in LDS
  int nslots;
  int lvl_begin[CL_MAX_LEVELS + 1];   // slots sorted by level
  const CoarseSlot* slots;
};
struct VCycleArgs {
  int nu, coarsest, smoother;
  float omega, ooa, pcoef;
};
struct CoarseLDS {
  float z[CL_MAX_TILES][512];
  float z2[CL_MAX_TILES][512];
  float b[CL_MAX_TILES][512];
  uint8_t ct[CL_MAX_TILES][512];
  CoarseSlot S[CL_MAX_TILES];
};

// A x and D of cell (i,j,k) of slot s with the subtree view at level sl.
__device__ __forceinline__ void cl_stencil(const CoarseLDS& M, const float (*Z)[512], int s, int i, int j, int k, int sl,
                                           float& ax, float& dg) {
  const CoarseSlot& S = M.S[s];
  const int c = (i << 6) | (j << 3) | k;
  const uint8_t to = subtree_type(S.type, S.level, sl);
  const float ooh = 1.0f / S.h;
  const float cll = ooh * ooh, clg = ((2.0f * ooh) / 3.0f) * ooh;
  const uint8_t c0 = M.ct[s][c];
  const float x0 = Z[s][c];
  const bool n0 = c0 & C_NEU;
  float sum = 0.f, d = 0.f;
  const int di[6] = {-1, 1, 0, 0, 0, 0}, dj[6] = {0, 0, -1, 1, 0, 0}, dk[6] = {0, 0, 0, 0, -1, 1};
  const int fi[6] = {0, 3, 1, 4, 2, 5};
#pragma unroll
  for (int e = 0; e < 6; e++) {
    int ni = i + di[e], nj = j + dj[e], nk = k + dk[e];
    float xv;
    uint8_t cv, tn;
    if (ni >= 0 && ni < 8 && nj >= 0 && nj < 8 && nk >= 0 && nk < 8) {
      const int nc = (ni << 6) | (nj << 3) | nk;
      xv = Z[s][nc];
      cv = M.ct[s][nc];
      tn = to;
    } else {
      const int q = S.nbr[fi[e]];
      if (q < 0) {
        xv = 0.f;
        cv = C_DIR;
        tn = to;
      } else {
        const int nc = ((ni & 7) << 6) | ((nj & 7) << 3) | (nk & 7);
        xv = Z[q][nc];
        cv = M.ct[q][nc];
        tn = subtree_type(M.S[q].type, M.S[q].level, sl);
      }
    }
    const float cf = (n0 || (cv & C_NEU)) ? 0.f : base_coeff(to, tn, cll, clg);
    sum = __builtin_fmaf(cf, x0 - xv, sum);
    d = d + cf;
  }
  const float h = S.h;
  const bool in = c0 & C_INT;
  ax = in ? ((sum * h) * h) * h : 0.f;
  dg = in ? ((d * h) * h) * h : 0.f;
}

__global__ __launch_bounds__(CL_THREADS) void k_coarse_lds(Ctx c, CoarsePlan cp, VCycleArgs v, float* Z, const float* R,
                                                           int flags) {
  __shared__ CoarseLDS M;
  if ((flags & F_CHECK_DONE) && c.ps[0].done) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int q = tid; q < cp.nslots; q += nt) M.S[q] = cp.slots[q];
  __syncthreads();
  // load cell types and the rhs; z starts at zero
  for (int idx = tid; idx < cp.nslots * 64; idx += nt) {
    const int s = idx >> 6, lane = idx & 63;
    const CoarseSlot& S = M.S[s];
    const size_t base = size_t(S.tile) * 512 + lane * 8;
    const uint64_t ct = load_ct8(c.ctype + base);
    *reinterpret_cast<uint64_t*>(&M.ct[s][lane * 8]) = ct;
    float b[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (S.type == T_LEAF) {
      load8(R + base, b);
      if (S.level == cp.C) {
        // ghost children below the LDS levels were restricted into gext
        const DevTile T = c.tiles[S.tile];
        if (T.gchild >= 0) {
          const int i = lane >> 3, j = lane & 7;
          for (int hh = 0; hh < 2; hh++) {
            int o = (i >> 2) | ((j >> 2) << 1) | (hh << 2);
            int gs = c.gcslot[T.gchild * 8 + o];
            if (gs < 0) continue;
            const float4 e = *reinterpret_cast<const float4*>(c.gext + size_t(gs) * 64 + (i & 3) * 16 + (j & 3) * 4);
            b[4 * hh + 0] = b[4 * hh + 0] + e.x;
            b[4 * hh + 1] = b[4 * hh + 1] + e.y;
            b[4 * hh + 2] = b[4 * hh + 2] + e.z;
            b[4 * hh + 3] = b[4 * hh + 3] + e.w;
          }
        }
      }
    } else if (S.type == T_NONLEAF && S.level == cp.C) {
      load8(c.rhsv + base, b);
    }
    for (int k = 0; k < 8; k++) {
      M.b[s][lane * 8 + k] = b[k];
      M.z[s][lane * 8 + k] = 0.f;
    }
  }
  __syncthreads();
  const bool jac = v.smoother != 0;
  // one red-black colour pass over the active slots of level l
  auto gs = [&](int l, int color) {
    const int s0 = cp.lvl_begin[l], ns = cp.lvl_begin[l + 1] - s0;
    for (int idx = tid; idx < ns * 256; idx += nt) {
      const int s = s0 + (idx >> 8), r = idx & 255;
      if (M.S[s].type == T_GHOST) continue;
      const int i = r >> 5, j = (r >> 2) & 7, k = 2 * (r & 3) + ((color + i + j) & 1);
      float ax, dg;
      cl_stencil(M, M.z, s, i, j, k, l, ax, dg);
      const int cc = (i << 6) | (j << 3) | k;
      float delta = 0.f;
      if ((M.ct[s][cc] & C_INT) && dg != 0.f) delta = (M.b[s][cc] - ax) / dg;
      M.z[s][cc] = M.z[s][cc] + delta;
    }
    __syncthreads();
  };
  // one damped-Jacobi sweep over the active slots of level l
  auto jacobi = [&](int l) {
    const int s0 = cp.lvl_begin[l], ns = cp.lvl_begin[l + 1] - s0;
    for (int idx = tid; idx < ns * 512; idx += nt) {
      const int s = s0 + (idx >> 9), cc = idx & 511;
      if (M.S[s].type == T_GHOST) continue;
      float ax, dg;
      cl_stencil(M, M.z, s, cc >> 6, (cc >> 3) & 7, cc & 7, l, ax, dg);
      float out = M.z[s][cc];
      if (M.ct[s][cc] & C_INT) out = (dg == 0.f) ? 0.f : M.z[s][cc] + (v.omega * (M.b[s][cc] - ax)) / dg;
      M.z2[s][cc] = out;
    }
    __syncthreads();
    for (int idx = tid; idx < ns * 512; idx += nt) {
      const int s = s0 + (idx >> 9), cc = idx & 511;
      if (M.S[s].type != T_GHOST) M.z[s][cc] = M.z2[s][cc];
    }
    __syncthreads();
  };
  auto smooth = [&](int l, int order) {
    for (int it = 0; it < v.nu; it++) {
      if (jac) jacobi(l);
      else { gs(l, order); gs(l, 1 - order); }
    }
  };
  for (int l = cp.C; l >= 1; l--) {
    smooth(l, 0);
    // residual of level l restricted to level l-1: one thread per coarse cell
    const int s0 = cp.lvl_begin[l], ns = cp.lvl_begin[l + 1] - s0;
    for (int idx = tid; idx < ns * 64; idx += nt) {
      const int s = s0 + (idx >> 6), q = idx & 63;
      const int a = q >> 4, bq = (q >> 2) & 3, cq = q & 3;
      float sum = 0.f;
      for (int ii = 0; ii < 2; ii++)
        for (int jj = 0; jj < 2; jj++)
          for (int kk = 0; kk < 2; kk++) {
            const int fi = 2 * a + ii, fj = 2 * bq + jj, fk = 2 * cq + kk;
            float ax, dg;
            cl_stencil(M, M.z, s, fi, fj, fk, l, ax, dg);
            const float rhs = (M.S[s].type == T_GHOST) ? 0.f : M.b[s][(fi << 6) | (fj << 3) | fk];
            sum += rhs - ax;
          }
      const CoarseSlot& S = M.S[s];
      const int ox = S.octant & 1, oy = (S.octant >> 1) & 1, oz = (S.octant >> 2) & 1;
      const int pc = ((ox * 4 + a) << 6) | ((oy * 4 + bq) << 3) | (oz * 4 + cq);
      if (S.type == T_GHOST) M.b[S.parent][pc] = M.b[S.parent][pc] + sum;
      else M.b[S.parent][pc] = sum * v.ooa;
    }
    __syncthreads();
  }
  const int it0 = v.coarsest / 2;
  for (int i = 0; i < it0; i++) { gs(0, 0); gs(0, 1); }
  for (int i = 0; i < it0; i++) { gs(0, 1); gs(0, 0); }
  for (int l = 1; l <= cp.C; l++) {
    const int s0 = cp.lvl_begin[l], ns = cp.lvl_begin[l + 1] - s0;
    for (int idx = tid; idx < ns * 512; idx += nt) {
      const int s = s0 + (idx >> 9), cc = idx & 511;
      const CoarseSlot& S = M.S[s];
      const int i = cc >> 6, j = (cc >> 3) & 7, k = cc & 7;
      const int ox = S.octant & 1, oy = (S.octant >> 1) & 1, oz = (S.octant >> 2) & 1;
      const float cz = M.z[S.parent][((ox * 4 + (i >> 1)) << 6) | ((oy * 4 + (j >> 1)) << 3) | (oz * 4 + (k >> 1))];
      const bool in = M.ct[s][cc] & C_INT;
      if (S.type == T_GHOST) M.z[s][cc] = in ? cz : 0.f;
      else M.z[s][cc] = in ? __builtin_fmaf(v.pcoef, cz, M.z[s][cc]) : 0.f;
    }
    __syncthreads();
    smooth(l, 1);
  }
  // write back the active tiles' z
  for (int idx = tid; idx < cp.nslots * 64; idx += nt) {
    const int s = idx >> 6, lane = idx & 63;
    if (M.S[s].type == T_GHOST) continue;
    float o[8];
    for (int k = 0; k < 8; k++) o[k] = M.z[s][lane * 8 + k];
    store8(Z + size_t(M.S[s].tile) * 512 + lane * 8, o);
  }
}

// Empty kernel, for measuring the launch floor.
__global__ void k_empty(int* p) {
  if (p && threadIdx.x == 1024) *p = 0;
}

// ---------------------------------------------------------------------------
// CG kernels (rotated loop: V-cycle, gamma/beta, p, A p + alpha, x/r + check)
// ---------------------------------------------------------------------------
// Init: x = 0, p = 0 on LEAF tiles; |b|^2; threshold; zero-rhs early exit
// (GMGSolver.cu:537-555).
__global__ __launch_bounds__(64) void k_cg_init(Ctx c, const int32_t* __restrict__ list, int n, int flags) {
  int w = work_index(n, flags & F_SWIZZLE);
  if (w >= n) return;
  const int t = list[w];
  const DevTile T = c.tiles[t];
  const int lane = lane_id();
  const size_t base = size_t(t) * 512 + lane * 8;
  float zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  store8(c.x + base, zero);
  store8(c.p + base, zero);
  float rv[8];
  load8(c.r + base, rv);
  double a = 0.0;
  for (int k = 0; k < 8; k++) a += double(rv[k]) * double(rv[k]);
  a = wave_sum(a);
  if (lane == 0) c.part[T.leaf_slot] = a;
}

// gamma = r.z over LEAF tiles; beta = gamma / gamma_old (0 on the first
// iteration, where p = z: GMGSolver.cu:564-568, 629-638).
__global__ __launch_bounds__(64) void k_cg_dot_rz(Ctx c, const int32_t* __restrict__ list, int n, int flags) {
  int w = work_index(n, flags & F_SWIZZLE);
  if (w >= n) return;
  const int t = list[w];
  const DevTile T = c.tiles[t];
  if (prob_done(c, T, flags)) return;
  const int lane = lane_id();
  const size_t base = size_t(t) * 512 + lane * 8;
  float rv[8], zv[8];
  load8(c.r + base, rv);
  load8(c.z + base, zv);
  double a = 0.0;
  for (int k = 0; k < 8; k++) a += double(rv[k]) * double(zv[k]);
  a = wave_sum(a);
  if (lane == 0) c.part[T.leaf_slot] = a;
}

// p = z + beta p on INTERIOR cells of LEAF tiles (GMGSolver.cu:642-653), then
// the new p copied into the tile's ghost children (their x is the parent's
// INTERIOR value, PoissonGrid.cu:585-605), ready for the operator.
__global__ __launch_bounds__(64) void k_cg_update_p(Ctx c, const int32_t* __restrict__ list, int n, int flags) {
  int w = work_index(n, flags & F_SWIZZLE);
  if (w >= n) return;
  const int t = list[w];
  const DevTile T = c.tiles[t];
  if (prob_done(c, T, flags)) return;
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t base = size_t(t) * 512 + lane * 8;
  const float beta = float(c.ps[T.prob].beta);
  float pv[8], zv[8];
  load8(c.p + base, pv);
  load8(c.z + base, zv);
  const uint64_t ct = load_ct8(c.ctype + base);
  for (int k = 0; k < 8; k++)
    if (ct_at(ct, k) & C_INT) pv[k] = __builtin_fmaf(beta, pv[k], zv[k]);
  store8(c.p + base, pv);
  if (T.gchild < 0) return;
  __shared__ float Lp[512];
  __shared__ uint8_t Lc[512];
  for (int k = 0; k < 8; k++) { Lp[lane * 8 + k] = pv[k]; Lc[lane * 8 + k] = ct_at(ct, k); }
  __syncthreads();
  for (int o = 0; o < 8; o++) {
    int g = c.gchild[T.gchild * 8 + o];
    if (g < 0) continue;
    int ox = o & 1, oy = (o >> 1) & 1, oz = (o >> 2) & 1;
    int src = ((ox * 4 + (i >> 1)) * 8 + (oy * 4 + (j >> 1))) * 8 + oz * 4;
    float gv[8];
    for (int k = 0; k < 8; k++) gv[k] = (Lc[src + (k >> 1)] & C_INT) ? Lp[src + (k >> 1)] : 0.f;
    store8(c.p + size_t(g) * 512 + lane * 8, gv);
  }
}

// A p on LEAF and GHOST tiles (subtree level -1).  LEAF tiles write A p and,
// when they have no ghost children, their p.Ap partial; GHOST tiles write the
// 2x2x2 sums of their rows for the parent (AccumulateToParents128 with all
// cell types, PoissonGrid.cu:664-710).  Channel mode (F_VCYCLE_OUT clear)
// writes ghost rows to AX instead and takes no dot.
__global__ __launch_bounds__(64) void k_apply(Ctx c, const int32_t* __restrict__ list, int n, const float* __restrict__ X,
                                              float* __restrict__ AX, int sl, int diag, int dot, int flags) {
  int w = work_index(n, flags & F_SWIZZLE);
  if (w >= n) return;
  const int t = list[w];
  const DevTile T = c.tiles[t];
  if (prob_done(c, T, flags)) return;
  __shared__ TileLDS L;
  __shared__ float gl[512];
  Stencil S;
  stage_tile(c, t, T, X, X, false, false, sl, L, S);
  wave_sync();
  float ax[8], dg[8];
  if (diag) apply_pencil<false, true>(L, S, dg, ax);
  else apply_pencil<true, false>(L, S, ax, dg);
  const int lane = lane_id();
  const size_t base = size_t(t) * 512 + lane * 8;
  if (T.type == T_GHOST && dot) {
    for (int k = 0; k < 8; k++) gl[lane * 8 + k] = (ct_at(S.ct, k) & (C_INT | C_DIR | C_NEU)) ? ax[k] : 0.f;
    __syncthreads();
    c.gsum[size_t(T.gslot) * 64 + lane] = octet_sum(gl, lane);
    return;
  }
  store8(AX + base, ax);
  if (!dot || T.type != T_LEAF || T.gchild >= 0) return;
  double a = 0.0;
  for (int k = 0; k < 8; k++) a += double(S.own[k]) * double(ax[k]);
  a = wave_sum(a);
  if (lane == 0) c.part[T.leaf_slot] = a;
}

// LEAF tiles with ghost children: A p += ghost row sums (the += of
// AccumulateToParents128), then the p.Ap partial; the last tile of the problem
// forms alpha = gamma / p.Ap (GMGSolver.cu:580-582).
__global__ __launch_bounds__(64) void k_ghost_gather(Ctx c, const int32_t* __restrict__ list, int n, int flags) {
  int w = work_index(n, flags & F_SWIZZLE);
  if (w >= n) return;
  const int t = list[w];
  const DevTile T = c.tiles[t];
  if (prob_done(c, T, flags)) return;
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t base = size_t(t) * 512 + lane * 8;
  float av[8], pv[8];
  load8(c.ap + base, av);
  load8(c.p + base, pv);
  for (int h = 0; h < 2; h++) {
    int o = (i >> 2) | ((j >> 2) << 1) | (h << 2);
    int gs = c.gcslot[T.gchild * 8 + o];
    if (gs < 0) continue;
    const float4 e = *reinterpret_cast<const float4*>(c.gsum + size_t(gs) * 64 + (i & 3) * 16 + (j & 3) * 4);
    av[4 * h + 0] = av[4 * h + 0] + e.x;
    av[4 * h + 1] = av[4 * h + 1] + e.y;
    av[4 * h + 2] = av[4 * h + 2] + e.z;
    av[4 * h + 3] = av[4 * h + 3] + e.w;
  }
  store8(c.ap + base, av);
  double a = 0.0;
  for (int k = 0; k < 8; k++) a += double(pv[k]) * double(av[k]);
  a = wave_sum(a);
  if (lane == 0) c.part[T.leaf_slot] = a;
}

// x += alpha p, r -= alpha A p on INTERIOR cells of LEAF tiles
// (GMGSolver.cu:587-609), |r|^2, and the convergence test of
// GMGSolver.cu:611-619 (or the iteration limit) setting the problem's done flag.
__global__ __launch_bounds__(64) void k_cg_update_xr(Ctx c, const int32_t* __restrict__ list, int n, int flags) {
  int w = work_index(n, flags & F_SWIZZLE);
  if (w >= n) return;
  const int t = list[w];
  const DevTile T = c.tiles[t];
  if (prob_done(c, T, flags)) return;
  const int lane = lane_id();
  const size_t base = size_t(t) * 512 + lane * 8;
  ProbState* ps = c.ps + T.prob;
  const float alpha = float(ps->alpha);
  float xv[8], rv[8], pv[8], av[8];
  load8(c.x + base, xv);
  load8(c.r + base, rv);
  load8(c.p + base, pv);
  load8(c.ap + base, av);
  const uint64_t ct = load_ct8(c.ctype + base);
  for (int k = 0; k < 8; k++) {
    if (ct_at(ct, k) & C_INT) {
      xv[k] = __builtin_fmaf(alpha, pv[k], xv[k]);
      rv[k] = __builtin_fmaf(-alpha, av[k], rv[k]);
    }
  }
  store8(c.x + base, xv);
  store8(c.r + base, rv);
  double a = 0.0;
  for (int k = 0; k < 8; k++) a += double(rv[k]) * double(rv[k]);
  a = wave_sum(a);
  if (lane == 0) c.part[T.leaf_slot] = a;
}

// Per-problem finalisation of a reduction over the LEAF-tile partials (one
// wavefront per problem, launched after the kernel that wrote the partials).
enum : int { FIN_BB = 0, FIN_RZ = 1, FIN_PAP = 2, FIN_RR = 3, FIN_MEAN = 4 };
__global__ __launch_bounds__(64) void k_finalize(Ctx c, int mode, int flags) {
  ProbState* ps = c.ps + blockIdx.x;
  if (mode != FIN_BB && (flags & F_CHECK_DONE) && ps->done) return;
  if (mode == FIN_MEAN && !ps->pure_neumann) return;
  const double s = sum_partials(c.part, ps->leaf_begin, ps->leaf_count);
  double q = 0.0;
  if (mode == FIN_MEAN) q = sum_partials(c.part2, ps->leaf_begin, ps->leaf_count);
  if (lane_id() != 0) return;
  switch (mode) {
    case FIN_BB: {  // GMGSolver.cu:542-554
      ps->bb = s;
      ps->rr = s;
      double thr = ps->tol2 * s;
      ps->thr = thr > 2.2250738585072014e-308 ? thr : 2.2250738585072014e-308;
      ps->iter = 0;
      ps->converged = (s == 0.0) ? 1 : 0;
      ps->done = (s == 0.0 || ps->max_iters <= 0) ? 1 : 0;
      ps->gamma = 0.0;
      break;
    }
    case FIN_RZ:  // gamma_old = gamma; gamma = r.z; beta = gamma / gamma_old (GMGSolver.cu:629-638)
      ps->gamma_old = ps->gamma;
      ps->gamma = s;
      ps->beta = (ps->iter == 0) ? 0.0 : s / ps->gamma_old;
      break;
    case FIN_PAP:  // alpha = gamma / p.Ap (GMGSolver.cu:580-582)
      ps->pAp = s;
      ps->alpha = ps->gamma / s;
      break;
    case FIN_RR: {  // GMGSolver.cu:611-619
      ps->rr = s;
      int it = ps->iter + 1;
      ps->iter = it;
      if (s < ps->thr) { ps->converged = 1; ps->done = 1; }
      else if (it >= ps->max_iters) ps->done = 1;
      break;
    }
    case FIN_MEAN:
      ps->mean = s / q;
      break;
  }
}

// Pure-Neumann re-centring (MeanAsync + ReCenterLeafVoxels,
// PoissonGrid.cu:231-325, PoissonSolver.cu:356-368): mean of r over all cells
// of LEAF tiles divided by the INTERIOR count, subtracted on INTERIOR cells.
__global__ __launch_bounds__(64) void k_mean(Ctx c, const int32_t* __restrict__ list, int n, const float* __restrict__ X, int flags) {
  int w = work_index(n, flags & F_SWIZZLE);
  if (w >= n) return;
  const int t = list[w];
  const DevTile T = c.tiles[t];
  if (prob_done(c, T, flags)) return;
  ProbState* ps = c.ps + T.prob;
  if (!ps->pure_neumann) return;
  const int lane = lane_id();
  const size_t base = size_t(t) * 512 + lane * 8;
  float v[8];
  load8(X + base, v);
  const uint64_t ct = load_ct8(c.ctype + base);
  double a = 0.0, q = 0.0;
  for (int k = 0; k < 8; k++) { a += double(v[k]); q += (ct_at(ct, k) & C_INT) ? 1.0 : 0.0; }
  a = wave_sum(a);
  q = wave_sum(q);
  if (lane == 0) { c.part[T.leaf_slot] = a; c.part2[T.leaf_slot] = q; }
}
__global__ __launch_bounds__(64) void k_subtract_mean(Ctx c, const int32_t* __restrict__ list, int n, float* __restrict__ X, int flags) {
  int w = work_index(n, flags & F_SWIZZLE);
  if (w >= n) return;
  const int t = list[w];
  const DevTile T = c.tiles[t];
  if (prob_done(c, T, flags)) return;
  ProbState* ps = c.ps + T.prob;
  if (!ps->pure_neumann) return;
  const double mean = ps->mean;
  const size_t base = size_t(t) * 512 + lane_id() * 8;
  float v[8];
  load8(X + base, v);
  const uint64_t ct = load_ct8(c.ctype + base);
  for (int k = 0; k < 8; k++) if (ct_at(ct, k) & C_INT) v[k] = float(double(v[k]) - mean);
  store8(X + base, v);
}

// ---------------------------------------------------------------------------
// Channel-level kernels for the operator API
// ---------------------------------------------------------------------------
// Ghost x from the parent's INTERIOR value (PoissonGrid.cu:585-605).
__global__ __launch_bounds__(64) void k_propagate_ghosts(Ctx c, const int32_t* __restrict__ list, int n,
                                                         const float* __restrict__ XC, float* __restrict__ XF) {
  int w = blockIdx.x;
  if (w >= n) return;
  const int t = list[w];
  const DevTile T = c.tiles[t];
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t base = size_t(t) * 512 + lane * 8;
  int ox = T.octant & 1, oy = (T.octant >> 1) & 1, oz = (T.octant >> 2) & 1;
  const size_t pb = size_t(T.parent) * 512 + ((ox * 4 + (i >> 1)) << 6) + ((oy * 4 + (j >> 1)) << 3) + oz * 4;
  const float4 pv = *reinterpret_cast<const float4*>(XC + pb);
  const uint32_t pc = *reinterpret_cast<const uint32_t*>(c.ctype + pb);
  const float cv[4] = {pv.x, pv.y, pv.z, pv.w};
  const uint64_t ct = load_ct8(c.ctype + base);
  float out[8];
  for (int k = 0; k < 8; k++) {
    bool pin = uint8_t(pc >> (8 * (k >> 1))) & C_INT;
    out[k] = ((ct_at(ct, k) & C_INT) && pin) ? cv[k >> 1] : 0.f;
  }
  store8(XF + base, out);
}
// Ghost rows summed into the parent channel (AccumulateToParents128 with
// coefficient 1, additive, all cell types; PoissonGrid.cu:664-719).
__global__ __launch_bounds__(64) void k_accumulate_ghosts(Ctx c, const int32_t* __restrict__ list, int n, float* __restrict__ AX) {
  int w = blockIdx.x;
  if (w >= n) return;
  const int t = list[w];
  const DevTile T = c.tiles[t];
  const int lane = lane_id();
  const size_t base = size_t(t) * 512 + lane * 8;
  __shared__ float gl[512];
  float v[8];
  load8(AX + base, v);
  const uint64_t ct = load_ct8(c.ctype + base);
  for (int k = 0; k < 8; k++) gl[lane * 8 + k] = (ct_at(ct, k) & (C_INT | C_DIR | C_NEU)) ? v[k] : 0.f;
  __syncthreads();
  float s = octet_sum(gl, lane);
  size_t pc = parent_cell(T, lane);
  AX[pc] = AX[pc] + s * 1.0f;
}
// Per-tile dot partials over all cells of the listed tiles, then one total.
__global__ __launch_bounds__(64) void k_dot_tiles(Ctx c, const int32_t* __restrict__ list, int n, const float* __restrict__ A,
                                                  const float* __restrict__ Bv, double* __restrict__ part) {
  int w = blockIdx.x;
  if (w >= n) return;
  const int t = list[w];
  const size_t base = size_t(t) * 512 + lane_id() * 8;
  float a[8], b[8];
  load8(A + base, a);
  load8(Bv + base, b);
  double s = 0.0;
  for (int k = 0; k < 8; k++) s += double(a[k]) * double(b[k]);
  s = wave_sum(s);
  if (lane_id() == 0) part[w] = s;
}
__global__ __launch_bounds__(64) void k_sum_partials(const double* __restrict__ part, int n, double* __restrict__ out) {
  double s = sum_partials(part, 0, n);
  if (lane_id() == 0) *out = s;
}

// Copy whole tiles of a list from one channel to another.
__global__ __launch_bounds__(64) void k_copy_tiles(const int32_t* __restrict__ list, int n, const float* __restrict__ src,
                                                   float* __restrict__ dst) {
  int w = blockIdx.x;
  if (w >= n) return;
  const size_t base = size_t(list[w]) * 512 + lane_id() * 8;
  float v[8];
  load8(src + base, v);
  store8(dst + base, v);
}
// Fill whole tiles of a list with a value.
__global__ __launch_bounds__(64) void k_fill_tiles(const int32_t* __restrict__ list, int n, float* __restrict__ dst, float val) {
  int w = blockIdx.x;
  if (w >= n) return;
  const size_t base = size_t(list[w]) * 512 + lane_id() * 8;
  float v[8] = {val, val, val, val, val, val, val, val};
  store8(dst + base, v);
}
// dampedJacobiSolve pieces (GMGSolver.cu:479-525) on LEAF tiles:
// R = B - R on INTERIOR cells; then X += omega R / D (X = 0 where D == 0).
__global__ __launch_bounds__(64) void k_leaf_residual(Ctx c, const int32_t* __restrict__ list, int n,
                                                      const float* __restrict__ B, float* __restrict__ R) {
  int w = blockIdx.x;
  if (w >= n) return;
  const size_t base = size_t(list[w]) * 512 + lane_id() * 8;
  float b[8], r[8];
  load8(B + base, b);
  load8(R + base, r);
  const uint64_t ct = load_ct8(c.ctype + base);
  for (int k = 0; k < 8; k++) if (ct_at(ct, k) & C_INT) r[k] = b[k] - r[k];
  store8(R + base, r);
}
__global__ __launch_bounds__(64) void k_leaf_jacobi_update(Ctx c, const int32_t* __restrict__ list, int n, float* __restrict__ X,
                                                           const float* __restrict__ R, const float* __restrict__ D, float omega) {
  int w = blockIdx.x;
  if (w >= n) return;
  const size_t base = size_t(list[w]) * 512 + lane_id() * 8;
  float x[8], r[8], d[8];
  load8(X + base, x);
  load8(R + base, r);
  load8(D + base, d);
  const uint64_t ct = load_ct8(c.ctype + base);
  for (int k = 0; k < 8; k++)
    if (ct_at(ct, k) & C_INT) x[k] = (d[k] == 0.f) ? 0.f : x[k] + (omega * r[k]) / d[k];
  store8(X + base, x);
}

}  // namespace cg
