// HIP kernels of the MGPCG path for gfx950 (device code, included by mgpcg.hip).
//
// Layout in HBM (structure of arrays, one problem or many in one pool):
//   channel c of tile t, cell (i,j,k):  ch[c][t*512 + (i<<6 | j<<3 | k)]  (FP32)
//   cell types:                        ctype[t*512 + ...]                (u8)
//   worklists:                          arrays of 176-byte TileRec (one per tile
//                                       and phase), so a workgroup finds its
//                                       tile, neighbours and children in one
//                                       scalar load
//   per-problem scalars:                ProbState
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
#include <type_traits>
#include <stdint.h>

namespace cg {

enum : uint8_t { T_LEAF = 1, T_NONLEAF = 2, T_GHOST = 4 };
enum : uint8_t { C_DIR = 1, C_INT = 2, C_NEU = 4 };
typedef __bf16 bf16;
typedef _Float16 f16;

struct alignas(16) TileRec {
  int32_t nbr[6];     // same-level face neighbours x-,y-,z-,x+,y+,z+ (-1: none)
  int32_t parent;     // -1 on level 0
  int32_t prob;       // problem index in the pool
  int32_t leaf_slot;  // LEAF: slot of its dot partial; -1 otherwise
  int32_t gslot;      // GHOST: slot in the 64-value ghost buffer; -1 otherwise
  int32_t self;       // the tile's index
  float h;            // cell size of the tile's level
  uint32_t meta;      // level | type << 8 | octant << 16 | has_gchild << 24 | all_interior << 25 |
                      // inert_candidate << 26 | all_open << 27
  uint32_t ntypes;    // type of face neighbour f in bits 4f..4f+3 (0: none)
  float bc[7];        // face coefficients of the smoothing view (subtree at the
                      // tile's own level): faces x-,y-,z-,x+,y+,z+, then inside
  int32_t nbr_parent[6];  // parent tile of face neighbour f (-1: no neighbour or level 0)
  int32_t gch[8];     // LEAF: ghost child per octant (-1: none)
  int32_t gcs[8];     // ... and its ghost slot
  int32_t pad1;
};
static_assert(sizeof(TileRec) == 176, "TileRec is 176 bytes");
// Fields packed in dwords so that a workgroup fetches them with scalar loads.
__host__ __device__ __forceinline__ int tr_level(const TileRec& T) { return int(T.meta & 0xff); }
__host__ __device__ __forceinline__ uint8_t tr_type(const TileRec& T) { return uint8_t((T.meta >> 8) & 0xff); }
__host__ __device__ __forceinline__ int tr_octant(const TileRec& T) { return int((T.meta >> 16) & 0xff); }
__host__ __device__ __forceinline__ bool tr_has_gchild(const TileRec& T) { return ((T.meta >> 24) & 1) != 0; }
// every one of the tile's 512 cells is INTERIOR (its cell types need not be read)
__host__ __device__ __forceinline__ bool tr_all_int(const TileRec& T) { return ((T.meta >> 25) & 1) != 0; }
// a LEAF tile without ghost children and without an INTERIOR cell: when its
// right-hand side is zero as well (decided per solve), every CG iteration
// leaves each of its values as it found them (see k_cg_init)
__host__ __device__ __forceinline__ bool tr_inert_cand(const TileRec& T) { return ((T.meta >> 26) & 1) != 0; }
// every cell INTERIOR and no face cut by a NEUMANN cell: each face mask byte is
// M_OPEN_INT (the masks need not be read)
__host__ __device__ __forceinline__ bool tr_all_open(const TileRec& T) { return ((T.meta >> 27) & 1) != 0; }
__host__ __device__ __forceinline__ uint8_t tr_ntype(const TileRec& T, int f) { return uint8_t((T.ntypes >> (4 * f)) & 0xf); }

struct alignas(16) ProbState {
  double gamma, gamma_old, alpha, beta;
  double bb, thr, rr, pAp;
  double mean, tol2;
  double thr_abs;   // floor of the stop threshold (|r|^2 < max(tol2 |b|^2, thr_abs)); 0: none
  double tol_abs2;  // FIN_IRBB: thr_abs = tol_abs2 * |b_outer|^2
  double bb_b;      // |b_outer|^2 of an FP64 refinement (the outer right-hand side)
  int32_t iter, done, converged, max_iters;
  uint32_t counter;
  int32_t leaf_begin, leaf_count, pure_neumann;
  int32_t check_stride;  // test convergence after iterations k*check_stride; <= 0: never
  int32_t scale_exp;     // float16 storage: the right-hand side is solved scaled by 2^scale_exp
};

// Channel pointers passed to every kernel by value.
struct Ctx {
  const uint8_t* ctype;
  const uint8_t* fmask;  // per cell: bit e (faces x-, x+, y-, y+, z-, z+) set if the face is not cut by a
                         // NEUMANN cell on either side, bit 6 set on INTERIOR cells (see k_face_mask)
  const uint32_t* fsplit;  // the same masks by colour: per tile [colour][lane], byte m = cell k = 2m + p
                           // of the lane's pencil, p = (colour + i + j) & 1 (k_face_mask)
  float* x;      // pressure / CG x
  float* r;      // rhs b, then residual r
  float* p;      // CG direction
  float* ap;     // A p
  float* z;      // preconditioned residual (V-cycle x)
  float* rhsv;   // V-cycle rhs: r on LEAF tiles (plus ghost sums), restricted residual on NONLEAF
  float* zb;     // damped-Jacobi ping-pong buffer; RB-GS down-sweep values
  bf16* zh;      // z, rhsv and zb stored as bfloat16 (mixed-precision form; null until used)
  bf16* rhsvh;
  bf16* zbh;
  f16* zf;       // the same buffers holding float16 values (one storage type per solve)
  f16* rhsvf;
  f16* zbf;
  float* gsum;   // 64 per ghost: 2x2x2 sums of the ghost rows of A p
  ProbState* ps;
  int32_t* dflag; // per-problem copy of ProbState::done, polled by the host
  const int32_t* pmap;  // finalize workgroup -> problem (null: identity)
  double* part;  // per-leaf-slot dot partials
  double* part2; // second partial stream (mean counts)
  const float* diag;  // per cell, colour-split order: the V-cycle colour sweep's diagonal D (k_diag)
  const TileRec* tiles;  // every tile's record by tile index (null where the handle has none)
  uint8_t* inert = nullptr;  // per tile: 1 while the current solve leaves the tile untouched (k_cg_init)
};

// The V-cycle vectors of the storage type V.
template <class V> __device__ __forceinline__ V* vc_z(const Ctx& c);
template <class V> __device__ __forceinline__ V* vc_rhs(const Ctx& c);
template <> __device__ __forceinline__ float* vc_z<float>(const Ctx& c) { return c.z; }
template <> __device__ __forceinline__ bf16* vc_z<bf16>(const Ctx& c) { return c.zh; }
template <> __device__ __forceinline__ float* vc_rhs<float>(const Ctx& c) { return c.rhsv; }
template <> __device__ __forceinline__ bf16* vc_rhs<bf16>(const Ctx& c) { return c.rhsvh; }
template <> __device__ __forceinline__ f16* vc_z<f16>(const Ctx& c) { return c.zf; }
template <> __device__ __forceinline__ f16* vc_rhs<f16>(const Ctx& c) { return c.rhsvf; }
template <class V> __device__ __forceinline__ V* vc_zb(const Ctx& c);
template <> __device__ __forceinline__ float* vc_zb<float>(const Ctx& c) { return c.zb; }
template <> __device__ __forceinline__ bf16* vc_zb<bf16>(const Ctx& c) { return c.zbh; }
template <> __device__ __forceinline__ f16* vc_zb<f16>(const Ctx& c) { return c.zbf; }

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__host__ __device__ __forceinline__ uint8_t subtree_type(uint8_t type, int level, int sl) {
  return (level == sl && type == T_NONLEAF) ? T_LEAF : type;
}
// Face coefficient before the NEUMANN cut (GMGSolver.cu:26-39): 1/h^2 on
// LEAF-LEAF faces, (2/3)/h^2 on LEAF-GHOST faces, 0 otherwise.
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
// V-cycle vectors may be stored as bfloat16 (8-bit significand, the FP32
// exponent range) and are always computed on in FP32: a stored value widens
// exactly, and every store rounds to nearest even (v_cvt_pk_bf16_f32).  One
// pencil of 8 values is then one 16-byte access.
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float to_f(float v) { return v; }
__device__ __forceinline__ float to_f(bf16 v) { return float(v); }
template <class V> __device__ __forceinline__ V from_f(float v);
template <> __device__ __forceinline__ float from_f<float>(float v) { return v; }
template <> __device__ __forceinline__ bf16 from_f<bf16>(float v) { return bf16(v); }
// the value a store of v to V-typed memory leaves (identity for FP32)
template <class V> __device__ __forceinline__ float vround(float v) { return to_f(from_f<V>(v)); }
__device__ __forceinline__ float bf_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
__device__ __forceinline__ uint32_t bf_pack(float a, float b) { return __builtin_bit_cast(uint32_t, bf16x2{bf16(a), bf16(b)}); }
__device__ __forceinline__ void load8(const bf16* __restrict__ p, float v[8]) {
  const uint4 q = *reinterpret_cast<const uint4*>(p);
  v[0] = bf_lo(q.x); v[1] = bf_hi(q.x); v[2] = bf_lo(q.y); v[3] = bf_hi(q.y);
  v[4] = bf_lo(q.z); v[5] = bf_hi(q.z); v[6] = bf_lo(q.w); v[7] = bf_hi(q.w);
}
__device__ __forceinline__ void store8(bf16* __restrict__ p, const float v[8]) {
  *reinterpret_cast<uint4*>(p) = make_uint4(bf_pack(v[0], v[1]), bf_pack(v[2], v[3]), bf_pack(v[4], v[5]), bf_pack(v[6], v[7]));
}
// a pencil as two float4 (k 0..3, 4..7)
__device__ __forceinline__ void load_pencil(const float* __restrict__ p, float4& a, float4& b) {
  const float4* q = reinterpret_cast<const float4*>(p);
  a = q[0];
  b = q[1];
}
__device__ __forceinline__ void load_pencil(const bf16* __restrict__ p, float4& a, float4& b) {
  const uint4 q = *reinterpret_cast<const uint4*>(p);
  a = make_float4(bf_lo(q.x), bf_hi(q.x), bf_lo(q.y), bf_hi(q.y));
  b = make_float4(bf_lo(q.z), bf_hi(q.z), bf_lo(q.w), bf_hi(q.w));
}
__device__ __forceinline__ void store_pencil(float* __restrict__ p, float4 a, float4 b) {
  float4* q = reinterpret_cast<float4*>(p);
  q[0] = a;
  q[1] = b;
}
__device__ __forceinline__ void store_pencil(bf16* __restrict__ p, float4 a, float4 b) {
  *reinterpret_cast<uint4*>(p) = make_uint4(bf_pack(a.x, a.y), bf_pack(a.z, a.w), bf_pack(b.x, b.y), bf_pack(b.z, b.w));
}
// four consecutive values (a parent quad)
__device__ __forceinline__ float4 load4(const float* __restrict__ p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ float4 load4(const bf16* __restrict__ p) {
  const uint2 q = *reinterpret_cast<const uint2*>(p);
  return make_float4(bf_lo(q.x), bf_hi(q.x), bf_lo(q.y), bf_hi(q.y));
}
__device__ __forceinline__ void store4(float* __restrict__ p, float4 a) { *reinterpret_cast<float4*>(p) = a; }
__device__ __forceinline__ void store4(bf16* __restrict__ p, float4 a) {
  *reinterpret_cast<uint2*>(p) = make_uint2(bf_pack(a.x, a.y), bf_pack(a.z, a.w));
}

// two consecutive values
__device__ __forceinline__ float2 load2(const float* __restrict__ p) { return *reinterpret_cast<const float2*>(p); }
__device__ __forceinline__ float2 load2(const bf16* __restrict__ p) {
  const uint32_t q = *reinterpret_cast<const uint32_t*>(p);
  return make_float2(bf_lo(q), bf_hi(q));
}
__device__ __forceinline__ void store2(float* __restrict__ p, float2 a) { *reinterpret_cast<float2*>(p) = a; }
__device__ __forceinline__ void store2(bf16* __restrict__ p, float2 a) { *reinterpret_cast<uint32_t*>(p) = bf_pack(a.x, a.y); }

// Or as float16 (11-bit significand, largest finite value 65504): a store
// clamps to +-65504 and rounds to nearest even (v_cvt_f16_f32); a solve in this
// storage scales its right-hand side by a power of two (k_finalize FIN_SCALE)
// so that the values sit well inside that range.
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float to_f(f16 v) { return float(v); }
template <> __device__ __forceinline__ f16 from_f<f16>(float v) {
  v = v > 65504.f ? 65504.f : (v < -65504.f ? -65504.f : v);
  return f16(v);
}
__device__ __forceinline__ float h_lo(uint32_t w) { return float(__builtin_bit_cast(f16, uint16_t(w & 0xffffu))); }
__device__ __forceinline__ float h_hi(uint32_t w) { return float(__builtin_bit_cast(f16, uint16_t(w >> 16))); }
__device__ __forceinline__ uint32_t h_pack(float a, float b) {
  return __builtin_bit_cast(uint32_t, f16x2{from_f<f16>(a), from_f<f16>(b)});
}
__device__ __forceinline__ void load8(const f16* __restrict__ p, float v[8]) {
  const uint4 q = *reinterpret_cast<const uint4*>(p);
  v[0] = h_lo(q.x); v[1] = h_hi(q.x); v[2] = h_lo(q.y); v[3] = h_hi(q.y);
  v[4] = h_lo(q.z); v[5] = h_hi(q.z); v[6] = h_lo(q.w); v[7] = h_hi(q.w);
}
__device__ __forceinline__ void store8(f16* __restrict__ p, const float v[8]) {
  *reinterpret_cast<uint4*>(p) = make_uint4(h_pack(v[0], v[1]), h_pack(v[2], v[3]), h_pack(v[4], v[5]), h_pack(v[6], v[7]));
}
__device__ __forceinline__ void load_pencil(const f16* __restrict__ p, float4& a, float4& b) {
  const uint4 q = *reinterpret_cast<const uint4*>(p);
  a = make_float4(h_lo(q.x), h_hi(q.x), h_lo(q.y), h_hi(q.y));
  b = make_float4(h_lo(q.z), h_hi(q.z), h_lo(q.w), h_hi(q.w));
}
__device__ __forceinline__ void store_pencil(f16* __restrict__ p, float4 a, float4 b) {
  *reinterpret_cast<uint4*>(p) = make_uint4(h_pack(a.x, a.y), h_pack(a.z, a.w), h_pack(b.x, b.y), h_pack(b.z, b.w));
}
__device__ __forceinline__ float4 load4(const f16* __restrict__ p) {
  const uint2 q = *reinterpret_cast<const uint2*>(p);
  return make_float4(h_lo(q.x), h_hi(q.x), h_lo(q.y), h_hi(q.y));
}
__device__ __forceinline__ void store4(f16* __restrict__ p, float4 a) {
  *reinterpret_cast<uint2*>(p) = make_uint2(h_pack(a.x, a.y), h_pack(a.z, a.w));
}
__device__ __forceinline__ float2 load2(const f16* __restrict__ p) {
  const uint32_t q = *reinterpret_cast<const uint32_t*>(p);
  return make_float2(h_lo(q), h_hi(q));
}
__device__ __forceinline__ void store2(f16* __restrict__ p, float2 a) { *reinterpret_cast<uint32_t*>(p) = h_pack(a.x, a.y); }

__device__ __forceinline__ uint64_t load_ct8(const uint8_t* __restrict__ p) {
  return *reinterpret_cast<const uint64_t*>(p);
}
// The lane's 8 cell types, or eight INTERIOR bytes without a load when the
// record says every cell of the tile is INTERIOR.
__device__ __forceinline__ uint64_t tile_ct8(const TileRec& T, const uint8_t* __restrict__ p) {
  return tr_all_int(T) ? 0x0101010101010101ull * C_INT : load_ct8(p);
}

// A wavefront-uniform value (keeps record fields in scalar registers).
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
// v[g] for a lane-dependent g in 0..3, as register selects
// One of four values by g in 0..3 (two selects on g's bits: no branches).
__device__ __forceinline__ int pick4(int g, int a, int b, int c, int d) {
  const int lo = (g & 1) ? b : a;
  const int hi = (g & 1) ? d : c;
  return (g & 2) ? hi : lo;
}

// Wavefront sum of a double in the fixed xor-butterfly order (masks 32..1);
// every lane ends with the same bits.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = v + __shfl_xor(v, m, 64);
  return v;
}

// Block -> worklist position.  Consecutive workgroups go round-robin to the 8
// XCDs; giving each XCD a contiguous (Morton-ordered) run of tiles keeps the
// face halos of a tile in the L2 of the XCD that reads them.
__device__ __forceinline__ int work_index(int xcd_swizzle) {
  int b = blockIdx.x;
  if (!xcd_swizzle) return b;
  int per = (int(gridDim.x) + 7) >> 3;
  return (b & 7) * per + (b >> 3);
}

// Each tile is worked by exactly one wavefront with its own LDS region, so
// staging only needs ordering inside the wavefront.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Kernel flags
// ---------------------------------------------------------------------------
enum : int {
  F_ZERO_X = 1,      // X is implicitly zero (first sweep)
  F_GHOST_ZERO = 2,  // ghost X is zero (V-cycle downstroke)
  F_VCYCLE = 4,      // V-cycle mode: rhs from rhsv; restriction into rhsv (ghosts +=)
  F_SWIZZLE = 16,    // XCD-aware tile order
  F_CHECK_DONE = 32, // skip tiles of converged problems
  F_DOT_RZ = 64,     // last colour pass of a level's up-sweep: also the r.z partial of each LEAF tile
  F_INERT = 128,     // a CG iteration's launch: tiles the solve's init found inert are skipped
};

__device__ __forceinline__ bool prob_done(const Ctx& c, const TileRec& T, int flags) {
  return (flags & F_CHECK_DONE) && c.ps[T.prob].done;
}

__device__ __forceinline__ void unpack8(float4 a, float4 b, float v[8]) {
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
// Cells of one parity of a k-pencil held as two float4 (k 0..3, 4..7).
__device__ __forceinline__ float4 even4(float4 a, float4 b) { return make_float4(a.x, a.z, b.x, b.z); }
__device__ __forceinline__ float4 odd4(float4 a, float4 b) { return make_float4(a.y, a.w, b.y, b.w); }
__device__ __forceinline__ float4 sel4(bool p, float4 if_false, float4 if_true) {
  return make_float4(p ? if_true.x : if_false.x, p ? if_true.y : if_false.y, p ? if_true.z : if_false.z,
                     p ? if_true.w : if_false.w);
}
__device__ __forceinline__ float comp4(float4 v, int m) { return m == 0 ? v.x : (m == 1 ? v.y : (m == 2 ? v.z : v.w)); }

// ---------------------------------------------------------------------------
// Colour-split layout of the V-cycle arrays (z on every level, its rhs and the
// down-sweep scratch).  Within a tile the 256 cells of colour 0 ((i+j+k) even)
// come first, then the 256 of colour 1; lane l = (i,j) keeps its four cells of
// each colour at [colour*256 + l*4, +4) in k order.  A red-black colour pass
// then reads the updated colour's rhs and writes its x as contiguous halves,
// and a neighbour's other-colour quad lines up component by component with the
// lane's own updated quad.  Channels seen by the caller and the CG vectors keep
// the natural order.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ int split_cell(int i, int j, int k) {
  return (((i + j + k) & 1) << 8) | (((i << 3) | j) << 2) | (k >> 1);
}
__host__ __device__ __forceinline__ int split_index(int cell) { return split_cell(cell >> 6, (cell >> 3) & 7, cell & 7); }
// colour of the even-k cells of pencil pl = (i,j)
__device__ __forceinline__ int even_colour(int pl) { return ((pl >> 3) + pl) & 1; }
// pencil pl of a tile (tile = the tile's first value) as natural float4 halves
template <class V>
__device__ __forceinline__ void load_pencil_split(const V* __restrict__ tile, int pl, float4& a, float4& b) {
  const bool q0 = even_colour(pl);
  const float4 h0 = load4(tile + pl * 4), h1 = load4(tile + 256 + pl * 4);
  const float4 E = sel4(q0, h0, h1), O = sel4(q0, h1, h0);
  a = make_float4(E.x, O.x, E.y, O.y);
  b = make_float4(E.z, O.z, E.w, O.w);
}
template <class V>
__device__ __forceinline__ void store_pencil_split(V* __restrict__ tile, int pl, float4 a, float4 b) {
  const bool q0 = even_colour(pl);
  const float4 E = even4(a, b), O = odd4(a, b);
  store4(tile + pl * 4, sel4(q0, E, O));
  store4(tile + 256 + pl * 4, sel4(q0, O, E));
}
template <class V>
__device__ __forceinline__ void load8_split(const V* __restrict__ tile, int pl, float v[8]) {
  float4 a, b;
  load_pencil_split(tile, pl, a, b);
  unpack8(a, b, v);
}
template <class V>
__device__ __forceinline__ void store8_split(V* __restrict__ tile, int pl, const float v[8]) {
  store_pencil_split(tile, pl, make_float4(v[0], v[1], v[2], v[3]), make_float4(v[4], v[5], v[6], v[7]));
}
// natural cells (i,j,K0..K0+3) of pencil pl, K0 in {0, 4}
template <class V>
__device__ __forceinline__ float4 load_quad_split(const V* __restrict__ tile, int pl, int K0) {
  const bool q0 = even_colour(pl);
  const float2 h0 = load2(tile + pl * 4 + (K0 >> 1)), h1 = load2(tile + 256 + pl * 4 + (K0 >> 1));
  const float2 E = q0 ? h1 : h0, O = q0 ? h0 : h1;
  return make_float4(E.x, O.x, E.y, O.y);
}
template <class V>
__device__ __forceinline__ void store_quad_split(V* __restrict__ tile, int pl, int K0, float4 v) {
  const bool q0 = even_colour(pl);
  const float2 E = make_float2(v.x, v.z), O = make_float2(v.y, v.w);
  store2(tile + pl * 4 + (K0 >> 1), q0 ? O : E);
  store2(tile + 256 + pl * 4 + (K0 >> 1), q0 ? E : O);
}

// Raw loads: a quad or a single value of V-typed storage as loaded (FP32
// values, or the bits of 16-bit ones), widened to FP32 only after every load of
// a tile has been issued; a widening inside the branch that loads a value would
// make that branch wait for its load.  hold() puts an empty asm the compiler
// cannot see through on the raw bits, so that it cannot move the widening up.
// A device array of zeros as large as one tile of FP64 values: loads of values
// that are 0 by definition read it (L2-resident) instead of being skipped.
// Nothing writes it; it is not declared const so that it sits in global
// memory like the arrays it stands in for (one kind of load for both).
__device__ __align__(16) float kZeroPage[1024] = {};
template <class V> __device__ __forceinline__ const V* zero_page() { return reinterpret_cast<const V*>(kZeroPage); }
// The face masks of a tile whose masks are all M_OPEN_INT (tr_all_open), in
// both layouts (512 cell bytes; 128 colour-split words), read in place of the
// tile's own: a load that stays in the caches instead of one from HBM.
#define CG_R8(x) x, x, x, x, x, x, x, x
__device__ __align__(16) uint32_t kOpenPage[128] = {CG_R8(CG_R8(0x7f7f7f7fu)), CG_R8(CG_R8(0x7f7f7f7fu))};
#undef CG_R8
__device__ __forceinline__ const uint8_t* own_masks(const Ctx& c, const TileRec& T) {
  return tr_all_open(T) ? reinterpret_cast<const uint8_t*>(kOpenPage) : c.fmask + size_t(T.self) * 512;
}
__device__ __forceinline__ const uint32_t* own_split_masks(const Ctx& c, const TileRec& T) {
  return tr_all_open(T) ? kOpenPage : c.fsplit + size_t(T.self) * 128;
}

template <class V> struct RawV;
template <> struct RawV<float> { typedef float4 Q; typedef float S; typedef float2 S2; };
template <> struct RawV<bf16> { typedef uint2 Q; typedef uint32_t S; typedef uint32_t S2; };
template <> struct RawV<f16> { typedef uint2 Q; typedef uint32_t S; typedef uint32_t S2; };
// a pair of consecutive values
__device__ __forceinline__ float2 load2_raw(const float* p) { return *reinterpret_cast<const float2*>(p); }
__device__ __forceinline__ uint32_t load2_raw(const bf16* p) { return *reinterpret_cast<const uint32_t*>(p); }
__device__ __forceinline__ uint32_t load2_raw(const f16* p) { return *reinterpret_cast<const uint32_t*>(p); }
template <class V> __device__ __forceinline__ float2 pair_f(typename RawV<V>::S2 v);
template <> __device__ __forceinline__ float2 pair_f<float>(float2 v) { return v; }
template <> __device__ __forceinline__ float2 pair_f<bf16>(uint32_t v) { return make_float2(bf_lo(v), bf_hi(v)); }
template <> __device__ __forceinline__ float2 pair_f<f16>(uint32_t v) { return make_float2(h_lo(v), h_hi(v)); }
__device__ __forceinline__ float4 load_q_raw(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ uint2 load_q_raw(const bf16* p) { return *reinterpret_cast<const uint2*>(p); }
__device__ __forceinline__ uint2 load_q_raw(const f16* p) { return *reinterpret_cast<const uint2*>(p); }
__device__ __forceinline__ float load_s_raw(const float* p) { return *p; }
__device__ __forceinline__ uint32_t load_s_raw(const bf16* p) { return *reinterpret_cast<const uint16_t*>(p); }
__device__ __forceinline__ uint32_t load_s_raw(const f16* p) { return *reinterpret_cast<const uint16_t*>(p); }
template <class V> __device__ __forceinline__ float4 quad_f(typename RawV<V>::Q q);
template <> __device__ __forceinline__ float4 quad_f<float>(float4 q) { return q; }
template <> __device__ __forceinline__ float4 quad_f<bf16>(uint2 q) { return make_float4(bf_lo(q.x), bf_hi(q.x), bf_lo(q.y), bf_hi(q.y)); }
template <> __device__ __forceinline__ float4 quad_f<f16>(uint2 q) { return make_float4(h_lo(q.x), h_hi(q.x), h_lo(q.y), h_hi(q.y)); }
template <class V> __device__ __forceinline__ float single_f(typename RawV<V>::S v);
template <> __device__ __forceinline__ float single_f<float>(float v) { return v; }
template <> __device__ __forceinline__ float single_f<bf16>(uint32_t v) { return bf_lo(v); }
template <> __device__ __forceinline__ float single_f<f16>(uint32_t v) { return h_lo(v); }
__device__ __forceinline__ void hold(float4&) {}
__device__ __forceinline__ void hold(float&) {}
// the same for an FP32 value whose later select must not pull its wait forward
__device__ __forceinline__ void hold_f(float& v) { __asm__("" : "+v"(v)); }
__device__ __forceinline__ void hold(uint2& q) { __asm__("" : "+v"(q.x), "+v"(q.y)); }
__device__ __forceinline__ void hold(uint32_t& v) { __asm__("" : "+v"(v)); }

// Raw loads of a colour-split pencil (both colour halves' quads), widened and
// put in natural order only after every load of the tile is out.
template <class V> struct RawPencil { typename RawV<V>::Q h0, h1; };
template <class V>
__device__ __forceinline__ RawPencil<V> load_pencil_split_raw(const V* __restrict__ tile, int pl) {
  RawPencil<V> r;
  r.h0 = load_q_raw(tile + pl * 4);
  r.h1 = load_q_raw(tile + 256 + pl * 4);
  return r;
}
// load_pencil_split's arrangement of the raw halves
template <class V>
__device__ __forceinline__ void pencil_split_f(const RawPencil<V>& r, int pl, float4& a, float4& b) {
  const bool q0 = even_colour(pl);
  const float4 h0 = quad_f<V>(r.h0), h1 = quad_f<V>(r.h1);
  const float4 E = sel4(q0, h0, h1), O = sel4(q0, h1, h0);
  a = make_float4(E.x, O.x, E.y, O.y);
  b = make_float4(E.z, O.z, E.w, O.w);
}
// hold for every storage type (FP32 values too: their rearrangement must not
// pull the wait forward either)
__device__ __forceinline__ void hold_all(float4& q) {
  hold_f(q.x);
  hold_f(q.y);
  hold_f(q.z);
  hold_f(q.w);
}
__device__ __forceinline__ void hold_all(uint2& q) { hold(q); }
__device__ __forceinline__ void hold_all(float& v) { hold_f(v); }
__device__ __forceinline__ void hold_all(uint32_t& v) { hold(v); }
__device__ __forceinline__ void hold_all(float2& v) {
  hold_f(v.x);
  hold_f(v.y);
}
template <class V> __device__ __forceinline__ void hold_all(RawPencil<V>& r) {
  hold_all(r.h0);
  hold_all(r.h1);
}

// ---------------------------------------------------------------------------
// Tile + halo staging
// ---------------------------------------------------------------------------
struct TileLDS {
  float4 xs[2][100];   // pencils (i,j), i,j in -1..8, index (i+1)*10 + (j+1); [0]: k 0..3, [1]: k 4..7
};

// Face-mask bits of a cell (Ctx::fmask).
enum : int { M_INT = 64 };
// A cell with all six faces uncut and INTERIOR.
constexpr uint32_t M_OPEN_INT = 0x7fu;

// The masks of a pencil's four colour cells k = p, 2+p, 4+p, 6+p (byte m = cell
// 2m + p) from its eight (byte k = cell k).
__device__ __forceinline__ uint32_t quad_mask(uint64_t m, bool p) {
  const uint64_t s = p ? (m >> 8) : m;
  return uint32_t(s & 0xffu) | uint32_t((s >> 8) & 0xff00u) | uint32_t((s >> 16) & 0xff0000u) |
         uint32_t((s >> 24) & 0xff000000u);
}

struct Stencil {
  float own[8];
  uint64_t m;          // own face masks, one byte per cell
  float xkm, xkp;      // k-halo values (z- neighbour cell (i,j,7), z+ neighbour cell (i,j,0))
  float b_int, b_f[6]; // base coefficients: inside the tile and across face f
  float h;
};

// Stage tile T's X values (zeros where the flags say so) and face masks.
//   zero_all:   treat X as 0 everywhere (first sweep from a zero guess)
//   ghost_zero: treat X of GHOST tiles as 0 (downstroke of the V-cycle)
//   bcv:        the 7 face coefficients of the view (faces x-..z+, inside)
// X holds active tiles, XG ghost tiles (they differ only for the damped-Jacobi
// ping-pong, where ghost values stay in the primary buffer).  SPLIT: X and XG
// are V-cycle arrays in the colour-split layout (the staged pencils are in
// natural order either way).
template <class V>
struct StageRaw {
  RawPencil<V> own, halo;
  typename RawV<V>::S km, kp;
  uint64_t m;
};
// A pencil's raw loads in either layout: colour-split (SPLIT, both halves'
// quads) or natural (cells 0..3 and 4..7).
template <bool SPLIT, class V>
__device__ __forceinline__ RawPencil<V> load_pencil_any_raw(const V* __restrict__ tile, int pl) {
  if (SPLIT) return load_pencil_split_raw(tile, pl);
  RawPencil<V> r;
  r.h0 = load_q_raw(tile + pl * 8);
  r.h1 = load_q_raw(tile + pl * 8 + 4);
  return r;
}
template <bool SPLIT, class V>
__device__ __forceinline__ void pencil_any_f(const RawPencil<V>& r, int pl, float4& a, float4& b) {
  if (SPLIT) {
    pencil_split_f(r, pl, a, b);
  } else {
    a = quad_f<V>(r.h0);
    b = quad_f<V>(r.h1);
  }
}
// The halo ring's lanes: lane g*8 + t (g < 4) fetches one neighbour pencil
//   g = 0: x- (pencil (7,t) of nbr 0), 1: x+ ((0,t) of nbr 3), 2: y- ((t,7) of nbr 1), 3: y+ ((t,0) of nbr 4)
// into LDS slot ring_dst (pencil index (i+1)*10 + (j+1) of the 10x10 frame).
__device__ __forceinline__ int ring_face(int g) { return pick4(g, 0, 3, 1, 4); }
__device__ __forceinline__ int ring_src(int g, int t) { return pick4(g, 56 + t, t, t * 8 + 7, t * 8); }
__device__ __forceinline__ int ring_dst(int g, int t) { return pick4(g, t + 1, 91 + t, (t + 1) * 10, (t + 1) * 10 + 9); }

// The loads of stage_tile, as raw bits: the own pencil, the face masks, the
// halo pencil of lanes 0..31 and the two k halos.  Every load is
// unconditional: a value that counts as 0 (a zero guess, outside the grid, a
// ghost in the down-sweep, lanes 32..63 of the ring) is read from the zero
// page, so no load sits in a branch whose merge would make the wavefront wait.
template <bool SPLIT, class V>
__device__ __forceinline__ StageRaw<V> stage_issue(const Ctx& c, const TileRec& T, const V* __restrict__ X,
                                                   const V* __restrict__ XG, bool zero_all, bool ghost_zero) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const bool own_ghost = tr_type(T) == T_GHOST;
  const V* zp = zero_page<V>();
  StageRaw<V> r;
  const bool own_zero = zero_all || (ghost_zero && own_ghost);
  r.own = load_pencil_any_raw<SPLIT>(own_zero ? zp : (own_ghost ? XG : X) + size_t(T.self) * 512, lane);
  r.m = load_ct8(own_masks(c, T) + lane * 8);
  {
    const int g = (lane >> 3) & 3, t = lane & 7;
    const int f = ring_face(g);
    const int n = pick4(g, uni(T.nbr[0]), uni(T.nbr[3]), uni(T.nbr[1]), uni(T.nbr[4]));
    const bool gh = tr_ntype(T, f) == T_GHOST;
    // else x = 0 (outside the grid: GMGSolver.cu:120-122)
    const bool ok = lane < 32 && n >= 0 && !(zero_all || (ghost_zero && gh));
    r.halo = load_pencil_any_raw<SPLIT>(ok ? (gh ? XG : X) + size_t(n) * 512 : zp, ok ? ring_src(g, t) : 0);
  }
  // k halos: cell (i,j,7) of the z- neighbour, (i,j,0) of the z+ neighbour
  const int o7 = SPLIT ? split_cell(i, j, 7) : lane * 8 + 7, o0 = SPLIT ? split_cell(i, j, 0) : lane * 8;
  const int nm = T.nbr[2], np = T.nbr[5];
  const bool ghm = tr_ntype(T, 2) == T_GHOST, ghp = tr_ntype(T, 5) == T_GHOST;
  const bool okm = nm >= 0 && !(zero_all || (ghost_zero && ghm));
  const bool okp = np >= 0 && !(zero_all || (ghost_zero && ghp));
  r.km = load_s_raw(okm ? (ghm ? XG : X) + size_t(nm) * 512 + o7 : zp);
  r.kp = load_s_raw(okp ? (ghp ? XG : X) + size_t(np) * 512 + o0 : zp);
  return r;
}
// The rest of stage_tile: one wait, then the values into LDS and the stencil.
template <bool SPLIT, class V>
__device__ __forceinline__ void stage_commit(StageRaw<V>& r, const TileRec& T, const float bcv[7], TileLDS& L,
                                             Stencil& S) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  hold_all(r.own);
  hold_all(r.halo);
  hold_all(r.km);
  hold_all(r.kp);
  float4 a, b;
  pencil_any_f<SPLIT>(r.own, lane, a, b);
  S.own[0] = a.x; S.own[1] = a.y; S.own[2] = a.z; S.own[3] = a.w;
  S.own[4] = b.x; S.own[5] = b.y; S.own[6] = b.z; S.own[7] = b.w;
  S.m = r.m;
  const int pc = (i + 1) * 10 + (j + 1);
  L.xs[0][pc] = a;
  L.xs[1][pc] = b;
  if (lane < 32) {
    const int g = lane >> 3, t = lane & 7;
    float4 v0, v1;
    pencil_any_f<SPLIT>(r.halo, ring_src(g, t), v0, v1);
    L.xs[0][ring_dst(g, t)] = v0;
    L.xs[1][ring_dst(g, t)] = v1;
  }
  S.xkm = single_f<V>(r.km);
  S.xkp = single_f<V>(r.kp);
  S.h = T.h;
  S.b_int = bcv[6];
  for (int f = 0; f < 6; f++) S.b_f[f] = bcv[f];
}

// Stage tile T's X values (zeros where the flags say so) and face masks.
//   zero_all:   treat X as 0 everywhere (first sweep from a zero guess)
//   ghost_zero: treat X of GHOST tiles as 0 (downstroke of the V-cycle)
//   bcv:        the 7 face coefficients of the view (faces x-..z+, inside)
// X holds active tiles, XG ghost tiles (they differ only for the damped-Jacobi
// ping-pong, where ghost values stay in the primary buffer).  SPLIT: X and XG
// are V-cycle arrays in the colour-split layout (the staged pencils are in
// natural order either way).
template <bool SPLIT, class V>
__device__ __forceinline__ void stage_tile(const Ctx& c, const TileRec& T, const V* __restrict__ X,
                                           const V* __restrict__ XG, bool zero_all, bool ghost_zero,
                                           const float bcv[7], TileLDS& L, Stencil& S) {
  StageRaw<V> r = stage_issue<SPLIT>(c, T, X, XG, zero_all, ghost_zero);
  stage_commit<SPLIT>(r, T, bcv, L, S);
}

// The CG operator's staging, where p is not stored on GHOST tiles: a GHOST
// cell's value is its parent cell's if that cell is INTERIOR, else 0 (the
// copy into ghost tiles the reference makes before the operator,
// PoissonGrid.cu:585-605), read from the parent (a LEAF tile one level up)
// wherever the tile itself or a face neighbour in its halo is a GHOST.  FP32,
// natural layout; every load unconditional as in stage_issue.
struct StagePG {
  float4 own0, own1, halo0, halo1;
  uint32_t own_ct, halo_ct, km_ct, kp_ct;
  float km, kp;
  uint64_t m;
  bool halo_gh;  // the lane's ring pencil is a GHOST's (from its parent)
};
// Cell of `parent` that holds cell (i, j, 2 k4 .. 2 k4 + 1) of its child with
// octant o (the pencil's four parent cells start at k4 = 0).
__device__ __forceinline__ size_t pg_cell(int parent, int o, int i, int j, int k4) {
  return size_t(parent) * 512 +
         size_t((((o & 1) * 4 + (i >> 1)) * 8 + (((o >> 1) & 1) * 4 + (j >> 1))) * 8 + ((o >> 2) & 1) * 4 + k4);
}
__device__ __forceinline__ StagePG stage_issue_pg(const Ctx& c, const TileRec& T, const float* __restrict__ X) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const bool own_ghost = tr_type(T) == T_GHOST;
  const int oct = tr_octant(T);
  const float* zp = zero_page<float>();
  const uint32_t* zw = reinterpret_cast<const uint32_t*>(kZeroPage);
  StagePG r;
  // own pencil, or its four parent cells and their types
  const size_t po = pg_cell(own_ghost ? T.parent : 0, oct, i, j, 0);
  const float* own = own_ghost ? X + po : X + size_t(T.self) * 512 + lane * 8;
  r.own0 = *reinterpret_cast<const float4*>(own);
  r.own1 = *reinterpret_cast<const float4*>(own_ghost ? zp : own + 4);
  r.own_ct = *(own_ghost ? reinterpret_cast<const uint32_t*>(c.ctype + po) : zw);
  r.m = load_ct8(own_masks(c, T) + lane * 8);
  {  // ring: lanes 0..31, one neighbour pencil each (x = 0 outside the grid)
    const int g = (lane >> 3) & 3, t = lane & 7;
    const int f = ring_face(g);
    const int n = pick4(g, uni(T.nbr[0]), uni(T.nbr[3]), uni(T.nbr[1]), uni(T.nbr[4]));
    const int npar = pick4(g, uni(T.nbr_parent[0]), uni(T.nbr_parent[3]), uni(T.nbr_parent[1]), uni(T.nbr_parent[4]));
    const bool ok = lane < 32 && n >= 0;
    const bool gh = ok && tr_ntype(T, f) == T_GHOST;
    const int sp = ring_src(g, t);                        // the neighbour's pencil si * 8 + sj
    const int no = oct ^ (g < 2 ? 1 : 2);                 // its octant: across x or y
    const size_t pn = pg_cell(gh ? npar : 0, no, sp >> 3, sp & 7, 0);
    const float* h = !ok ? zp : (gh ? X + pn : X + size_t(n) * 512 + sp * 8);
    r.halo0 = *reinterpret_cast<const float4*>(h);
    r.halo1 = *reinterpret_cast<const float4*>((ok && !gh) ? h + 4 : zp);
    r.halo_ct = *(gh ? reinterpret_cast<const uint32_t*>(c.ctype + pn) : zw);
    r.halo_gh = gh;
  }
  // k halos: cell (i,j,7) of the z- neighbour, (i,j,0) of the z+ neighbour
  const int nm = T.nbr[2], np = T.nbr[5];
  const bool ghm = nm >= 0 && tr_ntype(T, 2) == T_GHOST, ghp = np >= 0 && tr_ntype(T, 5) == T_GHOST;
  const size_t pm = pg_cell(ghm ? T.nbr_parent[2] : 0, oct ^ 4, i, j, 3);
  const size_t pp = pg_cell(ghp ? T.nbr_parent[5] : 0, oct ^ 4, i, j, 0);
  r.km = *(nm < 0 ? zp : (ghm ? X + pm : X + size_t(nm) * 512 + lane * 8 + 7));
  r.kp = *(np < 0 ? zp : (ghp ? X + pp : X + size_t(np) * 512 + lane * 8));
  r.km_ct = *(ghm ? c.ctype + pm : reinterpret_cast<const uint8_t*>(zw));
  r.kp_ct = *(ghp ? c.ctype + pp : reinterpret_cast<const uint8_t*>(zw));
  return r;
}
// A ghost pencil from its four parent cells: cells 2m and 2m + 1 take parent
// value m if that cell is INTERIOR, else 0.
__device__ __forceinline__ void pg_expand(float4 q, uint32_t ct, float4& a, float4& b) {
  const float v0 = (ct & C_INT) ? q.x : 0.f, v1 = ((ct >> 8) & C_INT) ? q.y : 0.f;
  const float v2 = ((ct >> 16) & C_INT) ? q.z : 0.f, v3 = ((ct >> 24) & C_INT) ? q.w : 0.f;
  a = make_float4(v0, v0, v1, v1);
  b = make_float4(v2, v2, v3, v3);
}
__device__ __forceinline__ void stage_commit_pg(StagePG& r, const TileRec& T, const float bcv[7], TileLDS& L,
                                                Stencil& S) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  hold_all(r.own0);
  hold_all(r.own1);
  hold_all(r.halo0);
  hold_all(r.halo1);
  hold_all(r.km);
  hold_all(r.kp);
  hold_all(r.own_ct);
  hold_all(r.halo_ct);
  hold_all(r.km_ct);
  hold_all(r.kp_ct);
  float4 a = r.own0, b = r.own1;
  if (tr_type(T) == T_GHOST) pg_expand(r.own0, r.own_ct, a, b);
  S.own[0] = a.x; S.own[1] = a.y; S.own[2] = a.z; S.own[3] = a.w;
  S.own[4] = b.x; S.own[5] = b.y; S.own[6] = b.z; S.own[7] = b.w;
  S.m = r.m;
  const int pc = (i + 1) * 10 + (j + 1);
  L.xs[0][pc] = a;
  L.xs[1][pc] = b;
  if (lane < 32) {
    const int g = lane >> 3, t = lane & 7;
    float4 v0 = r.halo0, v1 = r.halo1;
    if (r.halo_gh) pg_expand(r.halo0, r.halo_ct, v0, v1);
    L.xs[0][ring_dst(g, t)] = v0;
    L.xs[1][ring_dst(g, t)] = v1;
  }
  const bool ghm = T.nbr[2] >= 0 && tr_ntype(T, 2) == T_GHOST, ghp = T.nbr[5] >= 0 && tr_ntype(T, 5) == T_GHOST;
  S.xkm = (ghm && !(r.km_ct & C_INT)) ? 0.f : r.km;
  S.xkp = (ghp && !(r.kp_ct & C_INT)) ? 0.f : r.kp;
  S.h = T.h;
  S.b_int = bcv[6];
  for (int f = 0; f < 6; f++) S.b_f[f] = bcv[f];
}

// Face coefficients of tile T seen with the subtree view at level sl
// (GMGSolver.cu:26-39): 1/h^2 and (2/h/3)/h as the reference forms them; an
// absent neighbour counts as the tile's own type (GMGSolver.cu:118-122).
__host__ __device__ __forceinline__ void view_coeffs(const TileRec& T, int sl, float bcv[7]) {
  const float ooh = 1.0f / T.h;
  const float cll = ooh * ooh;
  const float clg = ((2.0f * ooh) / 3.0f) * ooh;
  const uint8_t to = subtree_type(tr_type(T), tr_level(T), sl);
  bcv[6] = base_coeff(to, to, cll, clg);
  for (int f = 0; f < 6; f++) {
    const uint8_t tn = (T.nbr[f] < 0) ? to : subtree_type(tr_ntype(T, f), tr_level(T), sl);
    bcv[f] = base_coeff(to, tn, cll, clg);
  }
}


// The 7-point row of one cell: faces visited in the reference's order x-, x+,
// y-, y+, z-, z+ (GMGSolver.cu:62-73); returns A x and D (times h^3, 0 off
// INTERIOR cells).
// The NEUMANN cuts and the INTERIOR test come from the cell's face mask m: a
// NEUMANN cell on either side of a face cuts it (coefficient 0).
// BITS (the colour sweeps' four cells): bit e of m as all ones or zero (one
// bitfield extract), so that the coefficient or +0 comes from one AND rather
// than a compare and a select; the empty asm keeps the compiler from turning
// the AND back into a select.  The same value either way; the eight-cell
// operator rows keep the select, which needs fewer registers there.
template <bool BITS = false>
__device__ __forceinline__ void cell_row_m(float x0, uint32_t m, const float xn[6], const float bc[6], float h, float& ax,
                                           float& dg) {
  float s = 0.f, d = 0.f;
#pragma unroll
  for (int e = 0; e < 6; e++) {
    float cf;
    if (BITS) {
      uint32_t open = uint32_t(int32_t(m << (31 - e)) >> 31);
      __asm__("" : "+v"(open));
      cf = __uint_as_float(__float_as_uint(bc[e]) & open);
    } else {
      cf = (m & (1u << e)) ? bc[e] : 0.f;
    }
    s = __builtin_fmaf(cf, x0 - xn[e], s);
    d = d + cf;
  }
  const bool in = m & M_INT;
  ax = in ? ((s * h) * h) * h : 0.f;
  dg = in ? ((d * h) * h) * h : 0.f;
}

// The four i/j-neighbour pencils of a lane (values and cell types).
struct NbrPencils {
  float4 im0, im1, ip0, ip1, jm0, jm1, jp0, jp1;
};
__device__ __forceinline__ NbrPencils lds_nbrs(const TileLDS& L) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const int pc = (i + 1) * 10 + (j + 1);
  NbrPencils N;
  N.im0 = L.xs[0][pc - 10]; N.im1 = L.xs[1][pc - 10];
  N.ip0 = L.xs[0][pc + 10]; N.ip1 = L.xs[1][pc + 10];
  N.jm0 = L.xs[0][pc - 1];  N.jm1 = L.xs[1][pc - 1];
  N.jp0 = L.xs[0][pc + 1];  N.jp1 = L.xs[1][pc + 1];
  return N;
}

// A x (and its diagonal) on all 8 cells of the lane, given its neighbour pencils.
template <bool WANT_AX, bool WANT_D>
__device__ __forceinline__ void apply_pencil(const NbrPencils& N, const Stencil& S, float ax[8], float dg[8]) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  float xim[8], xip[8], xjm[8], xjp[8];
  unpack8(N.im0, N.im1, xim);
  unpack8(N.ip0, N.ip1, xip);
  unpack8(N.jm0, N.jm1, xjm);
  unpack8(N.jp0, N.jp1, xjp);
  const float bim = (i == 0) ? S.b_f[0] : S.b_int;
  const float bip = (i == 7) ? S.b_f[3] : S.b_int;
  const float bjm = (j == 0) ? S.b_f[1] : S.b_int;
  const float bjp = (j == 7) ? S.b_f[4] : S.b_int;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const float xn[6] = {xim[k], xip[k], xjm[k], xjp[k], (k == 0) ? S.xkm : S.own[k - 1], (k == 7) ? S.xkp : S.own[k + 1]};
    const float bc[6] = {bim, bip, bjm, bjp, (k == 0) ? S.b_f[2] : S.b_int, (k == 7) ? S.b_f[5] : S.b_int};
    float a, d;
    cell_row_m(S.own[k], ct_at(S.m, k), xn, bc, S.h, a, d);
    if (WANT_AX) ax[k] = a;
    if (WANT_D) dg[k] = d;
  }
}


// One red-black colour update on the lane's four cells of that colour: cell
// (i,j,k) is updated when (i+j+k) % 2 == color, i.e. k = 2m + p with the lane's
// parity p.  The operands of those four rows are gathered with whole-vector
// selects on p (no per-lane register indexing), so only the four rows are formed.
__device__ __forceinline__ void gs_pencil(const NbrPencils& N, const Stencil& S, const float b[8], int color,
                                          float4& out0, float4& out1) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const bool p = ((color + i + j) & 1) != 0;
  auto pick = [&](float4 a, float4 c) { return sel4(p, even4(a, c), odd4(a, c)); };
  const float4 XIM = pick(N.im0, N.im1);
  const float4 XIP = pick(N.ip0, N.ip1);
  const float4 XJM = pick(N.jm0, N.jm1);
  const float4 XJP = pick(N.jp0, N.jp1);
  const float4 o0 = make_float4(S.own[0], S.own[1], S.own[2], S.own[3]);
  const float4 o1 = make_float4(S.own[4], S.own[5], S.own[6], S.own[7]);
  const float4 EV = even4(o0, o1), OD = odd4(o0, o1);
  const float4 X0 = sel4(p, EV, OD);
  // k-1 and k+1 neighbours of cell 2m+p
  const float4 XKM = sel4(p, make_float4(S.xkm, OD.x, OD.y, OD.z), EV);
  const float4 XKP = sel4(p, OD, make_float4(EV.y, EV.z, EV.w, S.xkp));
  const float4 B = sel4(p, make_float4(b[0], b[2], b[4], b[6]), make_float4(b[1], b[3], b[5], b[7]));
  const float bim = (i == 0) ? S.b_f[0] : S.b_int;
  const float bip = (i == 7) ? S.b_f[3] : S.b_int;
  const float bjm = (j == 0) ? S.b_f[1] : S.b_int;
  const float bjp = (j == 7) ? S.b_f[4] : S.b_int;
  const int sh = p ? 8 : 0;
  float nv[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const uint32_t mb = (m < 2 ? uint32_t(S.m) : uint32_t(S.m >> 32)) >> (16 * (m & 1) + sh);  // bits 0..6 used
    const float bkm = (!p && m == 0) ? S.b_f[2] : S.b_int;
    const float bkp = (p && m == 3) ? S.b_f[5] : S.b_int;
    const float x0 = comp4(X0, m);
    const float xn[6] = {comp4(XIM, m), comp4(XIP, m), comp4(XJM, m), comp4(XJP, m), comp4(XKM, m), comp4(XKP, m)};
    const float bc[6] = {bim, bip, bjm, bjp, bkm, bkp};
    float ax, dg;
    cell_row_m<true>(x0, mb, xn, bc, S.h, ax, dg);
    float delta = 0.f;
    if ((mb & M_INT) && dg != 0.f) delta = (comp4(B, m) - ax) / dg;
    nv[m] = x0 + delta;
  }
  // put the four new values back at their parity positions
  out0 = sel4(p, make_float4(nv[0], o0.y, nv[1], o0.w), make_float4(o0.x, nv[0], o0.z, nv[1]));
  out1 = sel4(p, make_float4(nv[2], o1.y, nv[3], o1.w), make_float4(o1.x, nv[2], o1.z, nv[3]));
}

// ---------------------------------------------------------------------------
// Per-tile bodies (one wavefront each)
// ---------------------------------------------------------------------------

// The four updated cells of a red-black colour pass, in split-layout quads:
// cell k = 2m + p of lane (i,j) (p = (color+i+j)&1) with value X0.m, rhs B.m,
// i/j neighbours XIM..XJP (component m: the neighbours' other-colour quads line
// up with the lane's updated quad), the lane's other-colour quad XO and the
// k-halo values xkm (cell (i,j,-1), used when p = 0) and xkp ((i,j,8), p = 1).
__device__ __forceinline__ float4 gs_quad(float4 X0, float4 XO, float4 XIM, float4 XIP, float4 XJM, float4 XJP,
                                          float xkm, float xkp, float4 B, uint32_t mq, const float bcv[7], float h,
                                          int i, int j, bool p) {
  // k-1 and k+1 neighbours of cell 2m+p
  const float4 XKM = sel4(p, make_float4(xkm, XO.x, XO.y, XO.z), XO);
  const float4 XKP = sel4(p, XO, make_float4(XO.y, XO.z, XO.w, xkp));
  const float bi = bcv[6];
  const float bim = (i == 0) ? bcv[0] : bi;
  const float bip = (i == 7) ? bcv[3] : bi;
  const float bjm = (j == 0) ? bcv[1] : bi;
  const float bjp = (j == 7) ? bcv[4] : bi;
  float nv[4];
  if (__all(mq == M_OPEN_INT * 0x01010101u)) {
    // Every updated cell of the wavefront is INTERIOR with six uncut faces:
    // each face coefficient is its base coefficient, so the diagonal sum
    // depends on the cell only through the k faces.  Its sums are formed once
    // per lane in the same order as cell_row_m forms them (faces x-, x+, y-,
    // y+, z-, z+; adding a zero is exact), so every value is the general path's.
    const float d4 = (((0.f + bim) + bip) + bjm) + bjp;
    const float dmid = (d4 + bi) + bi;
    const float dlo = (d4 + bcv[2]) + bi;   // cell k = 0 (p = 0, m = 0)
    const float dhi = (d4 + bi) + bcv[5];   // cell k = 7 (p = 1, m = 3)
    const float gmid = ((dmid * h) * h) * h;
    const float glo = ((dlo * h) * h) * h;
    const float ghi = ((dhi * h) * h) * h;
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const float bkm = (!p && m == 0) ? bcv[2] : bi;
      const float bkp = (p && m == 3) ? bcv[5] : bi;
      const float x0 = comp4(X0, m);
      float s = 0.f;
      s = __builtin_fmaf(bim, x0 - comp4(XIM, m), s);
      s = __builtin_fmaf(bip, x0 - comp4(XIP, m), s);
      s = __builtin_fmaf(bjm, x0 - comp4(XJM, m), s);
      s = __builtin_fmaf(bjp, x0 - comp4(XJP, m), s);
      s = __builtin_fmaf(bkm, x0 - comp4(XKM, m), s);
      s = __builtin_fmaf(bkp, x0 - comp4(XKP, m), s);
      const float ax = ((s * h) * h) * h;
      const float dg = (!p && m == 0) ? glo : ((p && m == 3) ? ghi : gmid);
      float delta = 0.f;
      if (dg != 0.f) delta = (comp4(B, m) - ax) / dg;
      nv[m] = x0 + delta;
    }
    return make_float4(nv[0], nv[1], nv[2], nv[3]);
  }
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const uint32_t mb = mq >> (8 * m);  // bits 0..6 used
    const float bkm = (!p && m == 0) ? bcv[2] : bi;
    const float bkp = (p && m == 3) ? bcv[5] : bi;
    const float x0 = comp4(X0, m);
    const float xn[6] = {comp4(XIM, m), comp4(XIP, m), comp4(XJM, m), comp4(XJP, m), comp4(XKM, m), comp4(XKP, m)};
    const float bc[6] = {bim, bip, bjm, bjp, bkm, bkp};
    float ax, dg;
    cell_row_m<true>(x0, mb, xn, bc, h, ax, dg);
    float delta = 0.f;
    if ((mb & M_INT) && dg != 0.f) delta = (comp4(B, m) - ax) / dg;
    nv[m] = x0 + delta;
  }
  return make_float4(nv[0], nv[1], nv[2], nv[3]);
}

struct GsLDS {
  float4 xo[100];  // other-colour quads of the tile's pencils and of the halo ring, index (i+1)*10 + (j+1)
};

// Red-black Gauss-Seidel colour sweep of a LEAF|NONLEAF tile (GMGSolver.cu:216-269)
// on the colour-split arrays: reads the tile's two x halves, the updated
// colour's rhs half and the neighbours' other-colour quads; writes the updated
// colour's half (and, from an implicit zero guess, zeros to the other half).
// An INTERIOR cell with D == 0 is left unchanged.
// The sweep is split into its loads (gs_issue: every global load of the tile,
// as raw bits into registers; gs_convert: to FP32), and the rest (gs_stage:
// the other-colour quads into LDS; gs_finish: the update and its stores), so
// that a wavefront issues every load of its tile (or tiles) before it waits
// on any.  A 16-bit value widens with one shift or conversion; done inside the
// branch that loads it, that would make the branch wait for its load, so the
// raw bits are held (an empty asm the compiler cannot see through) until all
// loads are out.
struct GsIn {
  float4 xc, xo, B, halo;
  float4 r0, r1;  // with F_DOT_RZ on a LEAF tile: r of the lane's pencil, cells 0..3 and 4..7
  uint32_t mask;  // the face masks of the lane's four cells of the pass's colour
  float xkm, xkp;
  int32_t done;   // the problem's done flag, loaded with the tile's data (used with F_CHECK_DONE only)
};
template <class V>
struct GsRaw {
  typename RawV<V>::Q xc, xo, B, halo;
  typename RawV<V>::S xkm, xkp;
  float4 r0, r1;
  uint32_t mask;
  int32_t done;
};
// DOT: the pass carries F_DOT_RZ (a separate instance, so that the other
// passes keep their register count).
template <class V, bool DOT = false>
__device__ __forceinline__ GsRaw<V> gs_issue(const Ctx& c, const TileRec& T, const V* X, const V* F, int color,
                                             int flags) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const bool zero_all = flags & F_ZERO_X, ghost_zero = flags & F_GHOST_ZERO;
  const int hc = color << 8, ho = (color ^ 1) << 8;
  // Every load below is unconditional: a value that is 0 by definition (x from
  // a zero guess, outside the grid, a ghost in the down-sweep, the halo of
  // lanes 32..63) is read from the zero page instead of skipped, so no load
  // sits in a branch whose merge would make the wavefront wait for it.
  const V* zp = zero_page<V>();
  const V* tile = zero_all ? zp : X + size_t(T.self) * 512;
  GsRaw<V> g;
  // one unconditional scalar load issued with the others, tested after the LDS
  // stage; without F_CHECK_DONE the face masks stand in as a readable address
  const int32_t* dp = (flags & F_CHECK_DONE) ? &c.ps[T.prob].done : reinterpret_cast<const int32_t*>(c.fmask);
  g.done = *dp;
  g.xc = load_q_raw(tile + hc + lane * 4);
  g.xo = load_q_raw(tile + ho + lane * 4);
  g.mask = own_split_masks(c, T)[color * 64 + lane];
  if (DOT) {
    // r of the lane's pencil on LEAF tiles (zeros otherwise)
    const float* rp = tr_type(T) == T_LEAF ? c.r + size_t(T.self) * 512 : zero_page<float>();
    const float4* r4 = reinterpret_cast<const float4*>(rp + lane * 8);
    g.r0 = r4[0];
    g.r1 = r4[1];
  }
  const V* src = (flags & F_VCYCLE) ? vc_rhs<V>(c) : F;
  g.B = load_q_raw(src + size_t(T.self) * 512 + hc + lane * 4);
  // halo ring: lanes 0..31 fetch the other-colour quad of one neighbour pencil
  //   q = 0: x- (pencil (7,t) of nbr 0), 1: x+ ((0,t) of nbr 3), 2: y- ((t,7) of nbr 1), 3: y+ ((t,0) of nbr 4)
  // x = 0 outside the grid (GMGSolver.cu:120-122)
  {
    const int q = (lane >> 3) & 3, t = lane & 7;
    const int f = pick4(q, 0, 3, 1, 4);
    const int n = pick4(q, uni(T.nbr[0]), uni(T.nbr[3]), uni(T.nbr[1]), uni(T.nbr[4]));
    const int src_pl = pick4(q, 56 + t, t, t * 8 + 7, t * 8);
    const bool gh = tr_ntype(T, f) == T_GHOST;
    const bool ok = lane < 32 && n >= 0 && !(zero_all || (ghost_zero && gh));
    g.halo = load_q_raw(ok ? X + size_t(n) * 512 + ho + src_pl * 4 : zp);
  }
  // k halos: cell (i,j,7) of the z- neighbour and (i,j,0) of the z+ neighbour
  {
    int n = T.nbr[2];
    bool ok = n >= 0 && !(zero_all || (ghost_zero && tr_ntype(T, 2) == T_GHOST));
    g.xkm = load_s_raw(ok ? X + size_t(n) * 512 + split_cell(i, j, 7) : zp);
    n = T.nbr[5];
    ok = n >= 0 && !(zero_all || (ghost_zero && tr_ntype(T, 5) == T_GHOST));
    g.xkp = load_s_raw(ok ? X + size_t(n) * 512 + split_cell(i, j, 0) : zp);
  }
  return g;
}
template <class V>
__device__ __forceinline__ GsIn gs_convert(GsRaw<V>& r) {
  hold(r.xc);
  hold(r.xo);
  hold(r.B);
  hold(r.halo);
  hold(r.xkm);
  hold(r.xkp);
  GsIn g;
  g.xc = quad_f<V>(r.xc);
  g.xo = quad_f<V>(r.xo);
  g.B = quad_f<V>(r.B);
  g.halo = quad_f<V>(r.halo);
  g.xkm = single_f<V>(r.xkm);
  g.xkp = single_f<V>(r.xkp);
  g.r0 = r.r0;
  g.r1 = r.r1;
  g.mask = r.mask;
  g.done = r.done;
  return g;
}
template <class V, bool DOT = false>
__device__ __forceinline__ GsIn gs_load(const Ctx& c, const TileRec& T, const V* X, const V* F, int color, int flags) {
  GsRaw<V> r = gs_issue<V, DOT>(c, T, X, F, color, flags);
  return gs_convert<V>(r);
}
__device__ __forceinline__ void gs_stage(const GsIn& g, GsLDS& L) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  L.xo[(i + 1) * 10 + (j + 1)] = g.xo;
  if (lane < 32) {
    const int q = lane >> 3, t = lane & 7;
    const int dst = (q == 0) ? (t + 1) : (q == 1) ? (91 + t) : (q == 2) ? ((t + 1) * 10) : ((t + 1) * 10 + 9);
    L.xo[dst] = g.halo;
  }
}
// after gs_stage and a wave_sync
template <class V, bool DOT = false>
__device__ __forceinline__ void gs_finish(const Ctx& c, const TileRec& T, V* X, int color, int flags, const GsIn& g,
                                          const GsLDS& L) {
  if ((flags & F_CHECK_DONE) && g.done) return;
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const int hc = color << 8, ho = (color ^ 1) << 8;
  V* tile = X + size_t(T.self) * 512;
  const int pc = (i + 1) * 10 + (j + 1);
  const bool p = ((color + i + j) & 1) != 0;
  const float4 nv = gs_quad(g.xc, g.xo, L.xo[pc - 10], L.xo[pc + 10], L.xo[pc - 1], L.xo[pc + 1], g.xkm, g.xkp, g.B,
                            g.mask, T.bc, T.h, i, j, p);
  store4(tile + hc + lane * 4, nv);
  if (flags & F_ZERO_X) store4(tile + ho + lane * 4, make_float4(0.f, 0.f, 0.f, 0.f));  // the zero guess of the other colour
  if (DOT && tr_type(T) == T_LEAF) {
    // this pass leaves the tile's z final: the updated colour as just stored,
    // the other colour as loaded; r.z over the lane's pencil in cell order,
    // as k_cg_dot_rz sums it
    const float zc[4] = {vround<V>(nv.x), vround<V>(nv.y), vround<V>(nv.z), vround<V>(nv.w)};
    const float zo[4] = {g.xo.x, g.xo.y, g.xo.z, g.xo.w};
    const float rv[8] = {g.r0.x, g.r0.y, g.r0.z, g.r0.w, g.r1.x, g.r1.y, g.r1.z, g.r1.w};
    double a = 0.0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const float z = ((k & 1) == int(p)) ? zc[k >> 1] : zo[k >> 1];
      a += double(rv[k]) * double(z);
    }
    a = wave_sum(a);
    if (lane == 0) c.part[T.leaf_slot] = a;
  }
}
template <class V, bool DOT>
__device__ __forceinline__ void tile_gs(const Ctx& c, const TileRec& T, int level, V* X, const V* F, int color,
                                        int flags, GsLDS& L) {
  const GsIn g = gs_load<V, DOT>(c, T, X, F, color, flags);
  gs_stage(g, L);
  wave_sync();
  gs_finish<V, DOT>(c, T, X, color, flags, g, L);
}

// stage_tile on the V-cycle arrays plus the tile's rhs b (natural order; 0 with
// zero_b), the rhs loaded with the stage's loads, ahead of the one wait.
// EAGER (the host's choice for small levels) stages the same way.
template <class V, bool EAGER = false>
__device__ __forceinline__ void stage_with_rhs(const Ctx& c, const TileRec& T, const V* X, const V* XG, const V* F,
                                               int flags, bool zero_b, TileLDS& L, Stencil& S, float b[8]) {
  const V* src = zero_b ? zero_page<V>() : ((flags & F_VCYCLE) ? vc_rhs<V>(c) : F) + size_t(T.self) * 512;
  StageRaw<V> r = stage_issue<true>(c, T, X, XG, flags & F_ZERO_X, flags & F_GHOST_ZERO);
  RawPencil<V> rb = load_pencil_split_raw(src, lane_id());
  hold_all(rb);
  stage_commit<true>(r, T, T.bc, L, S);
  float4 ba, bb;
  pencil_split_f(rb, lane_id(), ba, bb);
  b[0] = ba.x; b[1] = ba.y; b[2] = ba.z; b[3] = ba.w;
  b[4] = bb.x; b[5] = bb.y; b[6] = bb.z; b[7] = bb.w;
}

// Damped-Jacobi sweep (the update of GMGSolver.cu:508-525 with the level's
// operator): Y = X + omega (b - A X) / D on INTERIOR cells, 0 where D == 0.
template <class V>
__device__ __forceinline__ void tile_jacobi(const Ctx& c, const TileRec& T, int level, const V* X, const V* XG,
                                            V* Y, const V* F, float omega, int flags, TileLDS& L) {
  Stencil S;
  float b[8];
  stage_with_rhs(c, T, X, XG, F, flags, false, L, S, b);
  if (prob_done(c, T, flags)) return;
  wave_sync();
  float ax[8], dg[8];
  apply_pencil<true, true>(lds_nbrs(L), S, ax, dg);
  float out[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    out[k] = S.own[k];
    if (ct_at(S.m, k) & M_INT) {
      float res = b[k] - ax[k];
      out[k] = (dg[k] == 0.f) ? 0.f : S.own[k] + (omega * res) / dg[k];
    }
  }
  store8_split(Y + size_t(T.self) * 512, lane_id(), out);
}

// 2x2x2 sums of a tile's values staged in LDS, in the reference's (ii, jj, kk)
// order (GMGSolver.cu:319-328): lane q = (a,b,c) = (q>>4, (q>>2)&3, q&3) returns
// the sum of fine cells (2a+ii, 2b+jj, 2c+kk).
__device__ __forceinline__ float octet_sum(const float* Ls, int q) {
  int a = q >> 4, bq = (q >> 2) & 3, cq = q & 3;
  float s = 0.f;
  for (int ii = 0; ii < 2; ii++)
    for (int jj = 0; jj < 2; jj++)
      for (int kk = 0; kk < 2; kk++) s += Ls[((2 * a + ii) * 8 + (2 * bq + jj)) * 8 + 2 * cq + kk];
  return s;
}
__device__ __forceinline__ size_t parent_cell(const TileRec& T, int q) {
  int a = q >> 4, bq = (q >> 2) & 3, cq = q & 3;
  const int oct = tr_octant(T), ox = oct & 1, oy = (oct >> 1) & 1, oz = (oct >> 2) & 1;
  return size_t(T.parent) * 512 + ((ox * 4 + a) << 6) + ((oy * 4 + bq) << 3) + (oz * 4 + cq);
}
// the same parent cell in the colour-split layout
__device__ __forceinline__ size_t parent_cell_split(const TileRec& T, int q) {
  int a = q >> 4, bq = (q >> 2) & 3, cq = q & 3;
  const int oct = tr_octant(T), ox = oct & 1, oy = (oct >> 1) & 1, oz = (oct >> 2) & 1;
  return size_t(T.parent) * 512 + split_cell(ox * 4 + a, oy * 4 + bq, oz * 4 + cq);
}

// Residual and restriction of one tile of `level` to level-1
// (GMGSolver.cu:284-339).  Output channel CO (rhsv in V-cycle mode): the
// parent cell = sum * one_over_alpha, or += sum for a GHOST fine tile (whose
// LEAF parent's rhs already holds r).  In V-cycle mode a ghost's rhs is 0.
template <class V, bool EAGER = false>
__device__ __forceinline__ void tile_restrict(const Ctx& c, const TileRec& T, int level, const V* X, const V* F,
                                              V* CO, float one_over_alpha, int flags, TileLDS& L, float* res) {
  Stencil S;
  const int lane = lane_id();
  float b[8];
  stage_with_rhs<V, EAGER>(c, T, X, X, F, flags, (flags & F_VCYCLE) && tr_type(T) == T_GHOST, L, S, b);
  if (prob_done(c, T, flags)) return;
  wave_sync();
  float ax[8], dg[8];
  apply_pencil<true, false>(lds_nbrs(L), S, ax, dg);
  for (int k = 0; k < 8; k++) res[lane * 8 + k] = b[k] - ax[k];
  wave_sync();
  float s = octet_sum(res, lane);
  const bool ghost = tr_type(T) == T_GHOST;
  V* out = (flags & F_VCYCLE) ? vc_rhs<V>(c) : CO;
  size_t pcell = parent_cell_split(T, lane);
  if (ghost) out[pcell] = from_f<V>(to_f(out[pcell]) + s);
  else out[pcell] = from_f<V>(s * one_over_alpha);
}

// Prolongation and update of one tile of `level` (GMGSolver.cu:345-376):
// ghosts take the parent value, others add prolong_coeff * parent value, on
// INTERIOR cells; other cells become 0.  Ghost results go to XG (the primary
// buffer) and active tiles to XO (a Jacobi ping-pong target).
template <class V>
__device__ __forceinline__ void tile_prolong(const Ctx& c, const TileRec& T, int level, const V* XC, const V* XF,
                                             V* XO, V* XG, float pcoef, int flags) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t base = size_t(T.self) * 512 + lane * 8;
  const int oct = tr_octant(T), ox = oct & 1, oy = (oct >> 1) & 1, oz = (oct >> 2) & 1;
  const float4 pv = load_quad_split(XC + size_t(T.parent) * 512, (ox * 4 + (i >> 1)) * 8 + (oy * 4 + (j >> 1)), oz * 4);
  const float cz[4] = {pv.x, pv.y, pv.z, pv.w};
  const uint64_t ct = load_ct8(c.ctype + base);
  const bool ghost = tr_type(T) == T_GHOST;
  float fx[8];
  if (ghost || (flags & F_ZERO_X)) { for (int k = 0; k < 8; k++) fx[k] = 0.f; }
  else load8_split(XF + size_t(T.self) * 512, lane, fx);
  if (prob_done(c, T, flags)) return;
  float out[8];
  if (ghost) {
    for (int k = 0; k < 8; k++) out[k] = (ct_at(ct, k) & C_INT) ? cz[k >> 1] : 0.f;
    store8_split(XG + size_t(T.self) * 512, lane, out);
  } else {
    for (int k = 0; k < 8; k++) out[k] = (ct_at(ct, k) & C_INT) ? __builtin_fmaf(pcoef, cz[k >> 1], fx[k]) : 0.f;
    store8_split(XO + size_t(T.self) * 512, lane, out);
  }
}

// Prolongation fused into the first colour pass of the up-sweep on a level
// (GMGSolver.cu:345-376 then :216-269): every value the pass reads -- own cells
// and the face halos -- is formed on the fly as the prolongation would leave
// it (x + pcoef * parent on INTERIOR cells of active tiles, the parent value on
// INTERIOR cells of GHOST tiles, 0 elsewhere), so no separate launch is needed.
// XD holds the down-sweep values, XO receives the level's result; GHOST tiles
// in the list only store their prolongated values.
// parent cells of pencil (i,j) of a child in octant oct (colour-split array)
template <class V>
__device__ __forceinline__ float4 parent_quad(const V* XC, int parent, int oct, int i, int j) {
  const int ox = oct & 1, oy = (oct >> 1) & 1, oz = (oct >> 2) & 1;
  return load_quad_split(XC + size_t(parent) * 512, (ox * 4 + (i >> 1)) * 8 + (oy * 4 + (j >> 1)), oz * 4);
}
// prolongated values of one pencil: active (fma with the down-sweep values) or
// ghost; m: the pencil's face masks (their INTERIOR bits).  Values are rounded
// as a store to V-typed memory would leave them.
template <class V>
__device__ __forceinline__ void prolong_pencil(float4 xa, float4 xb, float4 pq, uint64_t m, bool ghost, float pcoef,
                                               float4& oa, float4& ob) {
  const float cz[4] = {pq.x, pq.y, pq.z, pq.w};
  float xo[8] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w};
  float o[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const bool in = ct_at(m, k) & M_INT;
    o[k] = in ? vround<V>(ghost ? cz[k >> 1] : __builtin_fmaf(pcoef, cz[k >> 1], xo[k])) : 0.f;
  }
  oa = make_float4(o[0], o[1], o[2], o[3]);
  ob = make_float4(o[4], o[5], o[6], o[7]);
}

// Prolongated k-halo value of lane (i,j): cell (i,j,KN) of face neighbour F.
template <int F, int KN, class V>
__device__ __forceinline__ void khalo_prolong(const Ctx& c, const TileRec& T, const V* XD, const V* XO, float pcoef,
                                              int oct, int lane, float& xv) {
  const int n = T.nbr[F];
  if (n < 0) { xv = 0.f; return; }
  const int i = lane >> 3, j = lane & 7;
  const uint8_t mv = c.fmask[size_t(n) * 512 + lane * 8 + KN];
  const bool gh = tr_ntype(T, F) == T_GHOST;
  const int noct = oct ^ 4;
  const int ox = noct & 1, oy = (noct >> 1) & 1, oz = (noct >> 2) & 1;
  const float pz = to_f(XO[size_t(T.nbr_parent[F]) * 512 + split_cell(ox * 4 + (i >> 1), oy * 4 + (j >> 1), oz * 4 + (KN >> 1))]);
  const float xo = gh ? 0.f : to_f(XD[size_t(n) * 512 + split_cell(i, j, KN)]);
  xv = (mv & M_INT) ? vround<V>(gh ? pz : __builtin_fmaf(pcoef, pz, xo)) : 0.f;
}

template <class V>
__device__ __forceinline__ void tile_gs_prolong(const Ctx& c, const TileRec& T, const V* XD, V* XO, float pcoef,
                                                int color, int flags, TileLDS& L) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const int oct = tr_octant(T);
  const bool own_ghost = tr_type(T) == T_GHOST;
  const uint64_t mown = load_ct8(own_masks(c, T) + lane * 8);
  const float4 pq = parent_quad(XO, T.parent, oct, i, j);
  float4 xa = make_float4(0.f, 0.f, 0.f, 0.f), xb = xa;
  if (!own_ghost) load_pencil_split(XD + size_t(T.self) * 512, lane, xa, xb);
  float4 oa, ob;
  prolong_pencil<V>(xa, xb, pq, mown, own_ghost, pcoef, oa, ob);
  if (own_ghost) {  // GHOST tile: its prolongated values only
    if (prob_done(c, T, flags)) return;
    store_pencil_split(XO + size_t(T.self) * 512, lane, oa, ob);
    return;
  }
  Stencil S;
  S.own[0] = oa.x; S.own[1] = oa.y; S.own[2] = oa.z; S.own[3] = oa.w;
  S.own[4] = ob.x; S.own[5] = ob.y; S.own[6] = ob.z; S.own[7] = ob.w;
  S.m = mown;
  const int pc = (i + 1) * 10 + (j + 1);
  L.xs[0][pc] = oa;
  L.xs[1][pc] = ob;
  // halo ring, lanes 0..31 (as stage_tile), prolongated in the neighbour's frame
  if (lane < 32) {
    const int g = lane >> 3, t = lane & 7;
    const int f = (g == 0) ? 0 : (g == 1) ? 3 : (g == 2) ? 1 : 4;
    // the record's fields as uniform values, so that the per-lane choice below is
    // a register select and not an indexed access to the record
    const int n = pick4(g, uni(T.nbr[0]), uni(T.nbr[3]), uni(T.nbr[1]), uni(T.nbr[4]));
    const int pn = pick4(g, uni(T.nbr_parent[0]), uni(T.nbr_parent[3]), uni(T.nbr_parent[1]), uni(T.nbr_parent[4]));
    const int ni = (g == 0) ? 7 : (g == 1) ? 0 : t, nj = (g == 2) ? 7 : (g == 3) ? 0 : t;  // pencil in the neighbour
    const int dst = (g == 0) ? (t + 1) : (g == 1) ? (91 + t) : (g == 2) ? ((t + 1) * 10) : ((t + 1) * 10 + 9);
    const int noct = oct ^ (1 << (f % 3));
    float4 va = make_float4(0.f, 0.f, 0.f, 0.f), vb = va;  // x = 0 outside the grid (GMGSolver.cu:120-122)
    if (n >= 0) {
      const bool gh = tr_ntype(T, f) == T_GHOST;
      const uint64_t w = load_ct8(c.fmask + size_t(n) * 512 + (ni * 8 + nj) * 8);
      float4 ya = make_float4(0.f, 0.f, 0.f, 0.f), yb = ya;
      if (!gh) load_pencil_split(XD + size_t(n) * 512, ni * 8 + nj, ya, yb);
      prolong_pencil<V>(ya, yb, parent_quad(XO, pn, noct, ni, nj), w, gh, pcoef, va, vb);
    }
    L.xs[0][dst] = va;
    L.xs[1][dst] = vb;
  }
  // k halos: cell (i,j,7) of the z- neighbour and (i,j,0) of the z+ neighbour
  khalo_prolong<2, 7>(c, T, XD, XO, pcoef, oct, lane, S.xkm);
  khalo_prolong<5, 0>(c, T, XD, XO, pcoef, oct, lane, S.xkp);
  S.h = T.h;
  S.b_int = T.bc[6];
  for (int f = 0; f < 6; f++) S.b_f[f] = T.bc[f];
  float b[8];
  load8_split(vc_rhs<V>(c) + size_t(T.self) * 512, lane, b);
  if (prob_done(c, T, flags)) return;
  wave_sync();
  float4 out0, out1;
  gs_pencil(lds_nbrs(L), S, b, color, out0, out1);
  store_pencil_split(XO + size_t(T.self) * 512, lane, out0, out1);
}

// ---------------------------------------------------------------------------
// One-tile-per-workgroup kernels over a worklist
// ---------------------------------------------------------------------------
#define TILE_KERNEL_PROLOGUE                     \
  int w = work_index(flags & F_SWIZZLE);         \
  if (w >= n) return;                            \
  const TileRec T = list[w];
// In a CG iteration's launch, a tile that k_cg_init found inert exits: the
// launch would write the values its arrays already hold.
#define SKIP_INERT \
  if ((flags & F_INERT) && tr_inert_cand(T) && c.inert[T.self]) return;

template <class V, bool DOT = false>
__global__ __launch_bounds__(64) void k_gs_color(Ctx c, const TileRec* __restrict__ list, int n, int level, V* X,
                                                 const V* F, int color, int flags) {
  TILE_KERNEL_PROLOGUE
  SKIP_INERT
  __shared__ GsLDS L;
  tile_gs<V, DOT>(c, T, level, X, F, color, flags, L);
}
// A colour sweep that writes the tile's whole result (the updated colour and
// the other colour as read) to Y and leaves X as it was, so that the next
// launch may read Y while it writes X.
template <class V>
__global__ __launch_bounds__(64) void k_gs_color_to(Ctx c, const TileRec* __restrict__ list, int n, const V* X, V* Y,
                                                    int color, int flags) {
  TILE_KERNEL_PROLOGUE
  __shared__ GsLDS L;
  const GsIn g = gs_load<V, false>(c, T, X, static_cast<const V*>(nullptr), color, flags);
  gs_stage(g, L);
  wave_sync();
  if ((flags & F_CHECK_DONE) && g.done) return;
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const int pc = (i + 1) * 10 + (j + 1);
  const bool p = ((color + i + j) & 1) != 0;
  const float4 nv = gs_quad(g.xc, g.xo, L.xo[pc - 10], L.xo[pc + 10], L.xo[pc - 1], L.xo[pc + 1], g.xkm, g.xkp, g.B,
                            g.mask, T.bc, T.h, i, j, p);
  V* tile = Y + size_t(T.self) * 512;
  store4(tile + (color << 8) + lane * 4, nv);
  store4(tile + ((color ^ 1) << 8) + lane * 4, g.xo);
}
// The same colour sweep with four wavefronts per workgroup, one tile each
// (tiles 4b .. 4b+3 of the list in workgroup b): a quarter of the workgroups
// to dispatch for the same tiles.
template <class V, bool DOT = false>
__global__ __launch_bounds__(256) void k_gs_color4(Ctx c, const TileRec* __restrict__ list, int n, int level, V* X,
                                                   const V* F, int color, int flags) {
  const int wave = threadIdx.x >> 6;
  const int w = work_index(flags & F_SWIZZLE) * 4 + wave;
  __shared__ GsLDS L[4];
  if (w >= n) return;  // whole wavefronts leave; the others synchronise within their wavefront only
  const TileRec T = list[w];
  tile_gs<V, DOT>(c, T, level, X, F, color, flags, L[wave]);
}
// The same colour sweep with a tile block of two: wavefront w works tiles 2w
// and 2w+1 of the list, loads of both in flight together.  Every tile's
// update is the one above (tiles of a colour pass are independent).
template <class V, bool DOT = false>
__global__ __launch_bounds__(64) void k_gs_color2(Ctx c, const TileRec* __restrict__ list, int n, int level, V* X,
                                                  const V* F, int color, int flags) {
  const int w = work_index(flags & F_SWIZZLE);
  const int t0 = 2 * w;
  if (t0 >= n) return;
  const bool two = t0 + 1 < n;
  const TileRec T0 = list[t0];
  const TileRec T1 = list[two ? t0 + 1 : t0];
  __shared__ GsLDS L[2];
  GsRaw<V> r0 = gs_issue<V, DOT>(c, T0, X, F, color, flags);
  GsRaw<V> r1{};  // unused without a second tile (no copy of r0, which would wait for its loads)
  if (two) r1 = gs_issue<V, DOT>(c, T1, X, F, color, flags);
  const GsIn g0 = gs_convert<V>(r0);
  const GsIn g1 = gs_convert<V>(r1);
  gs_stage(g0, L[0]);
  if (two) gs_stage(g1, L[1]);
  wave_sync();
  gs_finish<V, DOT>(c, T0, X, color, flags, g0, L[0]);
  if (two) gs_finish<V, DOT>(c, T1, X, color, flags, g1, L[1]);
}
// The same colour sweep with each wavefront walking its share of the list:
// while it updates one tile, the loads of its next tile are already out (a
// straight-line loop; after its last tile a wavefront loads that tile once
// more and drops it), so a wavefront's memory latency overlaps its own work
// instead of needing more wavefronts.  With F_SWIZZLE, the workgroups of XCD x
// (b % 8) walk the x-th eighth of the list (neighbouring tiles, one L2).  The
// grid is sized to the wavefronts the device holds at once.  Every tile's
// update is the one above.
template <class V, bool DOT = false>
__global__ __launch_bounds__(64) void k_gs_color_loop(Ctx c, const TileRec* __restrict__ list, int n, int level, V* X,
                                                      const V* F, int color, int flags) {
  __shared__ GsLDS L;
  const int b = blockIdx.x, nb = gridDim.x;
  int t, end, stride;
  if (flags & F_SWIZZLE) {
    const int per = (n + 7) >> 3, x = b & 7;
    t = x * per + (b >> 3);
    end = min(n, (x + 1) * per);
    stride = nb >> 3;  // nb is a multiple of 8
  } else {
    t = b;
    end = n;
    stride = nb;
  }
  if (t >= end) return;
  TileRec T = list[t];
  GsRaw<V> r = gs_issue<V, DOT>(c, T, X, F, color, flags);
  while (true) {
    const int tn = t + stride < end ? t + stride : t;
    const TileRec Tn = list[tn];
    GsRaw<V> rn = gs_issue<V, DOT>(c, Tn, X, F, color, flags);  // the next tile's loads
    const GsIn g = gs_convert<V>(r);
    gs_stage(g, L);
    wave_sync();
    gs_finish<V, DOT>(c, T, X, color, flags, g, L);
    if (tn == t) break;
    wave_sync();  // the LDS is staged again for the next tile
    t = tn;
    T = Tn;
    r = rn;
  }
}
// The colour sweep's diagonal of every cell of a tile (the D of gs_quad: the
// same coefficient choice and the same operations, in the colour-split order),
// for the fused first iteration below.  Fixed by the face masks and the tile's
// coefficients, so formed once when the problem is built.
__global__ __launch_bounds__(64) void k_diag(Ctx c, const TileRec* __restrict__ list, int n, float* __restrict__ D) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const uint64_t mask = load_ct8(c.fmask + size_t(T.self) * 512 + lane * 8);
  const float bi = T.bc[6];
  const float bim = (i == 0) ? T.bc[0] : bi;
  const float bip = (i == 7) ? T.bc[3] : bi;
  const float bjm = (j == 0) ? T.bc[1] : bi;
  const float bjp = (j == 7) ? T.bc[4] : bi;
  const float xn[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int col = 0; col < 2; col++) {
    const bool p = ((col + i + j) & 1) != 0;
    const int sh = p ? 8 : 0;
    float d[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const uint32_t mb = (m < 2 ? uint32_t(mask) : uint32_t(mask >> 32)) >> (16 * (m & 1) + sh);
      const float bkm = (!p && m == 0) ? T.bc[2] : bi;
      const float bkp = (p && m == 3) ? T.bc[5] : bi;
      const float bc[6] = {bim, bip, bjm, bjp, bkm, bkp};
      float ax;
      cell_row_m<true>(0.f, mb, xn, bc, T.h, ax, d[m]);
    }
    *reinterpret_cast<float4*>(D + size_t(T.self) * 512 + (col << 8) + lane * 4) = make_float4(d[0], d[1], d[2], d[3]);
  }
}

// The value a colour-0 cell holds after the first pass from the zero guess:
// 0 + (b - A 0) / D on INTERIOR cells with D != 0 (A 0 = +0), 0 elsewhere
// (D is 0 off INTERIOR cells), as stored in V.
template <class V>
__device__ __forceinline__ float first_red(float b, float d) {
  const float delta = (d != 0.f) ? (b - 0.f) / d : 0.f;
  return vround<V>(0.f + delta);
}

// The first red-black iteration of a down-sweep level from the zero guess in
// one launch: both colour passes of GMGSolver.cu:216-269 (colour 0 from x = 0,
// the other colour written 0; then colour 1).  After the first pass a colour-0
// cell depends on its own rhs and diagonal only, so the colour-0 values of the
// halo are formed from the neighbours' rhs and stored diagonals and the second
// pass needs no grid-wide wait.  Every value is the one the two passes give.
template <class V>
__global__ __launch_bounds__(64) void k_gs_first(Ctx c, const TileRec* __restrict__ list, int n, V* X, int flags) {
  TILE_KERNEL_PROLOGUE
  SKIP_INERT
  __shared__ GsLDS L;
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const bool ghost_zero = flags & F_GHOST_ZERO;
  const V* R = vc_rhs<V>(c);
  // ---- loads, all issued before the first wait
  const int32_t* dp = (flags & F_CHECK_DONE) ? &c.ps[T.prob].done : reinterpret_cast<const int32_t*>(c.fmask);
  const int32_t done = *dp;
  const uint64_t mask = load_ct8(own_masks(c, T) + lane * 8);
  typename RawV<V>::Q rb0 = load_q_raw(R + size_t(T.self) * 512 + lane * 4);        // colour-0 rhs quad
  typename RawV<V>::Q rb1 = load_q_raw(R + size_t(T.self) * 512 + 256 + lane * 4);  // colour-1 rhs quad
  // halo ring of the colour-1 pass: lanes 0..31, the colour-0 quad of one neighbour pencil
  // (a missing or ghost neighbour is read at the tile itself and its value
  // taken as 0 afterwards, so that the loads need no inner branch)
  typename RawV<V>::Q hb{};
  float4 hd = make_float4(0.f, 0.f, 0.f, 0.f);
  int dst = 0;
  bool hok = false;
  if (lane < 32) {
    const int q = lane >> 3, t = lane & 7;
    const int f = (q == 0) ? 0 : (q == 1) ? 3 : (q == 2) ? 1 : 4;
    const int nb = pick4(q, uni(T.nbr[0]), uni(T.nbr[3]), uni(T.nbr[1]), uni(T.nbr[4]));
    const int src_pl = (q == 0) ? (56 + t) : (q == 1) ? t : (q == 2) ? (t * 8 + 7) : (t * 8);
    dst = (q == 0) ? (t + 1) : (q == 1) ? (91 + t) : (q == 2) ? ((t + 1) * 10) : ((t + 1) * 10 + 9);
    const bool gh = tr_ntype(T, f) == T_GHOST;
    hok = nb >= 0 && !(ghost_zero && gh);  // x = 0 outside the grid and on ghosts in the down-sweep
    const int ln = hok ? nb : T.self;
    hb = load_q_raw(R + size_t(ln) * 512 + src_pl * 4);
    hd = *reinterpret_cast<const float4*>(c.diag + size_t(ln) * 512 + src_pl * 4);
  }
  // k halos: cell (i,j,7) of the z- neighbour and (i,j,0) of the z+ neighbour;
  // their value after the first pass is 0 unless they are colour-0 cells
  const int nzm = T.nbr[2], nzp = T.nbr[5];
  const bool kokm = nzm >= 0 && !(ghost_zero && tr_ntype(T, 2) == T_GHOST) && ((i + j + 7) & 1) == 0;
  const bool kokp = nzp >= 0 && !(ghost_zero && tr_ntype(T, 5) == T_GHOST) && ((i + j) & 1) == 0;
  const int lzm = nzm >= 0 ? nzm : T.self, lzp = nzp >= 0 ? nzp : T.self;
  typename RawV<V>::S kbm = load_s_raw(R + size_t(lzm) * 512 + split_cell(i, j, 7));
  typename RawV<V>::S kbp = load_s_raw(R + size_t(lzp) * 512 + split_cell(i, j, 0));
  float kdm = c.diag[size_t(lzm) * 512 + split_cell(i, j, 7)];
  float kdp = c.diag[size_t(lzp) * 512 + split_cell(i, j, 0)];
  hold(rb0);
  hold(rb1);
  hold(hb);
  hold_f(hd.x);
  hold_f(hd.y);
  hold_f(hd.z);
  hold_f(hd.w);
  hold(kbm);
  hold(kbp);
  hold_f(kdm);
  hold_f(kdp);
  if (!hok) hd = make_float4(0.f, 0.f, 0.f, 0.f);
  if (!kokm) kdm = 0.f;
  if (!kokp) kdp = 0.f;
  // ---- pass 1: the tile's colour-0 cells from x = 0 (gs_quad as the colour-0 pass runs it)
  const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
  const bool p0 = ((i + j) & 1) != 0, p1 = !p0;
  const float4 n0 = gs_quad(z4, z4, z4, z4, z4, z4, 0.f, 0.f, quad_f<V>(rb0), quad_mask(mask, p0), T.bc, T.h, i, j, p0);
  const float4 xr = make_float4(vround<V>(n0.x), vround<V>(n0.y), vround<V>(n0.z), vround<V>(n0.w));
  // the halo's colour-0 values as that pass leaves them
  const float4 hbf = quad_f<V>(hb);
  const float4 hv = make_float4(first_red<V>(hbf.x, hd.x), first_red<V>(hbf.y, hd.y), first_red<V>(hbf.z, hd.z),
                                first_red<V>(hbf.w, hd.w));
  const float xkm = first_red<V>(single_f<V>(kbm), kdm);
  const float xkp = first_red<V>(single_f<V>(kbp), kdp);
  L.xo[(i + 1) * 10 + (j + 1)] = xr;
  if (lane < 32) L.xo[dst] = hv;
  wave_sync();
  if ((flags & F_CHECK_DONE) && done) return;
  // ---- pass 2: colour 1, its own values 0 before the update
  const int pc = (i + 1) * 10 + (j + 1);
  const float4 n1 = gs_quad(z4, xr, L.xo[pc - 10], L.xo[pc + 10], L.xo[pc - 1], L.xo[pc + 1], xkm, xkp,
                            quad_f<V>(rb1), quad_mask(mask, p1), T.bc, T.h, i, j, p1);
  V* tile = X + size_t(T.self) * 512;
  store4(tile + lane * 4, xr);
  store4(tile + 256 + lane * 4, n1);
}

#include "gs_pair.hpp"

// The same pass with every global load issued first, as raw bits (own pencil,
// parent quad, face masks, halo pencils and their parents' quads, k halos, rhs,
// done flag) and one wait before the values are formed.  Missing neighbours
// and lanes 32..63 read the tile itself and its parent (values dropped), so no
// load sits in a branch.  Faster on small levels, where each launch waits on
// memory latency; slower on big ones (more loads and registers), so the host
// picks it by level size.  Every value is the one tile_gs_prolong forms.
template <class V>
struct KHaloRaw {
  uint32_t mv;
  typename RawV<V>::S pz, xo;
};
// The raw loads of one prolongated k-halo value: face mask, parent value and
// down-sweep value of cell (i,j,KN) of face neighbour F, each from the zero
// page where it does not exist (no neighbour, a ghost neighbour's down-sweep
// value, a ghost tile's own stage).
template <int F, int KN, class V>
__device__ __forceinline__ KHaloRaw<V> khalo_prolong_issue(const Ctx& c, const TileRec& T, const V* XD, const V* XO,
                                                           int oct, int lane, bool skip) {
  KHaloRaw<V> k;
  const int n = T.nbr[F];
  const bool gh = tr_ntype(T, F) == T_GHOST;
  const bool ok = !skip && n >= 0;
  const V* zp = zero_page<V>();
  const int i = lane >> 3, j = lane & 7;
  k.mv = (ok ? c.fmask + size_t(n) * 512 + lane * 8 : reinterpret_cast<const uint8_t*>(zp))[ok ? KN : 0];
  const int noct = oct ^ 4;
  const int ox = noct & 1, oy = (noct >> 1) & 1, oz = (noct >> 2) & 1;
  k.pz = load_s_raw(ok ? XO + size_t(T.nbr_parent[F]) * 512 +
                             split_cell(ox * 4 + (i >> 1), oy * 4 + (j >> 1), oz * 4 + (KN >> 1))
                       : zp);
  k.xo = load_s_raw(ok && !gh ? XD + size_t(n) * 512 + split_cell(i, j, KN) : zp);
  return k;
}
// khalo_prolong's value from the raw loads (a missing neighbour's zero mask
// gives 0)
template <int F, class V>
__device__ __forceinline__ float khalo_prolong_value(const TileRec& T, KHaloRaw<V> k, float pcoef) {
  const bool gh = tr_ntype(T, F) == T_GHOST;
  const float pz = single_f<V>(k.pz);
  const float xo = single_f<V>(k.xo);
  return (k.mv & M_INT) ? vround<V>(gh ? pz : __builtin_fmaf(pcoef, pz, xo)) : 0.f;
}
template <class V> struct RawQuadS { typename RawV<V>::S2 h0, h1; };
template <class V>
__device__ __forceinline__ RawQuadS<V> load_quad_split_raw(const V* __restrict__ tile, int pl, int K0) {
  RawQuadS<V> r;
  r.h0 = load2_raw(tile + pl * 4 + (K0 >> 1));
  r.h1 = load2_raw(tile + 256 + pl * 4 + (K0 >> 1));
  return r;
}
// load_quad_split's arrangement of the raw pairs
template <class V>
__device__ __forceinline__ float4 quad_split_f(const RawQuadS<V>& r, int pl) {
  const bool q0 = even_colour(pl);
  const float2 h0 = pair_f<V>(r.h0), h1 = pair_f<V>(r.h1);
  const float2 E = q0 ? h1 : h0, O = q0 ? h0 : h1;
  return make_float4(E.x, O.x, E.y, O.y);
}
template <class V>
__device__ __forceinline__ void tile_gs_prolong_eager(const Ctx& c, const TileRec& T, const V* XD, V* XO, float pcoef,
                                                      int color, int flags, TileLDS& L) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const int oct = tr_octant(T);
  const bool own_ghost = tr_type(T) == T_GHOST;
  const V* zp = zero_page<V>();
  // ---- loads, none in a branch (values that do not exist come from the zero page)
  const int32_t* dp = (flags & F_CHECK_DONE) ? &c.ps[T.prob].done : reinterpret_cast<const int32_t*>(c.fmask);
  const int32_t done = *dp;
  const uint64_t mown = load_ct8(own_masks(c, T) + lane * 8);
  const int ppl = ((oct & 1) * 4 + (i >> 1)) * 8 + (((oct >> 1) & 1) * 4 + (j >> 1));  // parent pencil
  RawQuadS<V> rpq = load_quad_split_raw(XO + size_t(T.parent) * 512, ppl, ((oct >> 2) & 1) * 4);
  RawPencil<V> rown = load_pencil_split_raw(own_ghost ? zp : XD + size_t(T.self) * 512, lane);
  RawPencil<V> rrhs = load_pencil_split_raw(own_ghost ? zp : vc_rhs<V>(c) + size_t(T.self) * 512, lane);
  // halo ring of lanes 0..31 as in tile_gs_prolong; every lane loads
  const int g = (lane >> 3) & 3, t = lane & 7;
  const int f = ring_face(g);
  const int hn = pick4(g, uni(T.nbr[0]), uni(T.nbr[3]), uni(T.nbr[1]), uni(T.nbr[4]));
  const int pn = pick4(g, uni(T.nbr_parent[0]), uni(T.nbr_parent[3]), uni(T.nbr_parent[1]), uni(T.nbr_parent[4]));
  const int ni = pick4(g, 7, 0, t, t), nj = pick4(g, t, t, 7, 0);  // pencil in the neighbour
  const int noct = oct ^ (1 << (f % 3));
  const int hpl = ni * 8 + nj;
  const int hppl = ((noct & 1) * 4 + (ni >> 1)) * 8 + (((noct >> 1) & 1) * 4 + (nj >> 1));
  const bool hgh = tr_ntype(T, f) == T_GHOST;
  const bool hok = !own_ghost && lane < 32 && hn >= 0;
  const uint64_t hw = load_ct8(hok ? c.fmask + size_t(hn) * 512 + hpl * 8 : reinterpret_cast<const uint8_t*>(zp));
  RawPencil<V> rh = load_pencil_split_raw(hok && !hgh ? XD + size_t(hn) * 512 : zp, hpl);
  RawQuadS<V> rhpq = load_quad_split_raw(hok ? XO + size_t(pn) * 512 : zp, hppl, ((noct >> 2) & 1) * 4);
  KHaloRaw<V> km = khalo_prolong_issue<2, 7>(c, T, XD, XO, oct, lane, own_ghost);
  KHaloRaw<V> kp = khalo_prolong_issue<5, 0>(c, T, XD, XO, oct, lane, own_ghost);
  // ---- one wait, then the values
  hold_all(rpq.h0);
  hold_all(rpq.h1);
  hold_all(rown);
  hold_all(rrhs);
  hold_all(rh);
  hold_all(rhpq.h0);
  hold_all(rhpq.h1);
  hold_all(km.mv);
  hold_all(km.pz);
  hold_all(km.xo);
  hold_all(kp.mv);
  hold_all(kp.pz);
  hold_all(kp.xo);
  float4 xa, xb;
  pencil_split_f(rown, lane, xa, xb);
  float4 oa, ob;
  prolong_pencil<V>(xa, xb, quad_split_f(rpq, ppl), mown, own_ghost, pcoef, oa, ob);
  if (own_ghost) {  // GHOST tile: its prolongated values only
    if ((flags & F_CHECK_DONE) && done) return;
    store_pencil_split(XO + size_t(T.self) * 512, lane, oa, ob);
    return;
  }
  Stencil S;
  S.own[0] = oa.x; S.own[1] = oa.y; S.own[2] = oa.z; S.own[3] = oa.w;
  S.own[4] = ob.x; S.own[5] = ob.y; S.own[6] = ob.z; S.own[7] = ob.w;
  S.m = mown;
  const int pc = (i + 1) * 10 + (j + 1);
  L.xs[0][pc] = oa;
  L.xs[1][pc] = ob;
  if (lane < 32) {
    // x = 0 outside the grid (GMGSolver.cu:120-122): a missing neighbour's
    // zero masks make every value 0
    float4 ya, yb, va, vb;
    pencil_split_f(rh, hpl, ya, yb);
    prolong_pencil<V>(ya, yb, quad_split_f(rhpq, hppl), hw, hgh, pcoef, va, vb);
    L.xs[0][ring_dst(g, t)] = va;
    L.xs[1][ring_dst(g, t)] = vb;
  }
  S.xkm = khalo_prolong_value<2>(T, km, pcoef);
  S.xkp = khalo_prolong_value<5>(T, kp, pcoef);
  S.h = T.h;
  S.b_int = T.bc[6];
  for (int e = 0; e < 6; e++) S.b_f[e] = T.bc[e];
  float4 ba, bb;
  pencil_split_f(rrhs, lane, ba, bb);
  const float b[8] = {ba.x, ba.y, ba.z, ba.w, bb.x, bb.y, bb.z, bb.w};
  if ((flags & F_CHECK_DONE) && done) return;
  wave_sync();
  float4 out0, out1;
  gs_pencil(lds_nbrs(L), S, b, color, out0, out1);
  store_pencil_split(XO + size_t(T.self) * 512, lane, out0, out1);
}
template <class V>
__global__ __launch_bounds__(64) void k_gs_prolong_eager(Ctx c, const TileRec* __restrict__ list, int n, const V* XD,
                                                         V* XO, float pcoef, int color, int flags) {
  TILE_KERNEL_PROLOGUE
  SKIP_INERT
  __shared__ TileLDS L;
  tile_gs_prolong_eager(c, T, XD, XO, pcoef, color, flags, L);
}
template <class V>
__global__ __launch_bounds__(64) void k_gs_prolong(Ctx c, const TileRec* __restrict__ list, int n, const V* XD,
                                                   V* XO, float pcoef, int color, int flags) {
  TILE_KERNEL_PROLOGUE
  SKIP_INERT
  __shared__ TileLDS L;
  tile_gs_prolong(c, T, XD, XO, pcoef, color, flags, L);
}
template <class V>
__global__ __launch_bounds__(64) void k_jacobi(Ctx c, const TileRec* __restrict__ list, int n, int level, const V* X,
                                               const V* XG, V* Y, const V* F, float omega, int flags) {
  TILE_KERNEL_PROLOGUE
  SKIP_INERT
  __shared__ TileLDS L;
  tile_jacobi(c, T, level, X, XG, Y, F, omega, flags, L);
}
template <class V, bool EAGER = false>
__global__ __launch_bounds__(64) void k_restrict(Ctx c, const TileRec* __restrict__ list, int n, int level, const V* X,
                                                 const V* F, V* CO, float one_over_alpha, int flags) {
  TILE_KERNEL_PROLOGUE
  SKIP_INERT
  __shared__ TileLDS L;
  __shared__ float res[512];
  tile_restrict<V, EAGER>(c, T, level, X, F, CO, one_over_alpha, flags, L, res);
}
template <class V>
__global__ __launch_bounds__(64) void k_prolong(Ctx c, const TileRec* __restrict__ list, int n, int level,
                                                const V* XC, const V* XF, V* XO, V* XG, float pcoef, int flags) {
  TILE_KERNEL_PROLOGUE
  SKIP_INERT
  tile_prolong(c, T, level, XC, XF, XO, XG, pcoef, flags);
}

// ---------------------------------------------------------------------------
// Coarse levels 0..C of the V-cycle (smoothing, restriction, the bottom solve
// on level 0, prolongation) for one problem in one workgroup, with every tile
// of those levels (active and ghost, at most CL_MAX_TILES) resident in LDS and
// one thread per cell.  The host picks C so that these levels are tiny; one
// launch then replaces the ~10 launches per level a grid-wide phase each would
// need.  The per-cell arithmetic is the same as the tile kernels', so results
// are identical; the bottom is GMGSolver.cu:432-433.
// ---------------------------------------------------------------------------
constexpr int CL_MAX_TILES = 12;     // tiles on levels 0..C together
constexpr int CL_MAX_PER_LEVEL = 8;  // tiles on any one of those levels
constexpr int CL_MAX_LEVELS = 8;
constexpr int CL_THREADS = 1024;
constexpr int CL_CPT = CL_MAX_PER_LEVEL * 256 / CL_THREADS;  // cells of one colour per thread
constexpr int CL_ZERO = CL_MAX_TILES * 512;                 // LDS index of a cell that stays 0
struct alignas(16) CoarseSlot {
  int32_t tile;      // tile index
  int32_t nbr[6];    // slot of the same-level face neighbour (x-,y-,z-,x+,y+,z+), -1: none
  int32_t parent;    // slot of the parent, -1 on level 0
  int32_t level;
  float h;
  uint8_t type, octant, pad[2];
  float bc[7];       // base coefficients at the slot's own level: faces x-,y-,z-,x+,y+,z+ and inside
};
static_assert(sizeof(CoarseSlot) == 72 || sizeof(CoarseSlot) == 80, "CoarseSlot layout");
struct CoarsePlan {
  int prob;                           // problem index in the pool
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
  float z[CL_MAX_TILES * 512 + 4];  // V-cycle x of every slot (slot * 512 + cell), then CL_ZERO = 0
  float b[CL_MAX_TILES][512];       // V-cycle rhs
  float t[CL_MAX_TILES][512];       // residuals (restriction) or new values (damped Jacobi)
  alignas(16) uint8_t ct[CL_MAX_TILES][512];
  CoarseSlot S[CL_MAX_TILES];
};

// The static part of one cell's row, prepared once per level phase and kept in
// registers: LDS indices of the cell and its six neighbours (faces x-, x+, y-,
// y+, z-, z+; CL_ZERO outside the grid), the face coefficients after the
// NEUMANN cut, the diagonal and the rhs.  A pass then costs seven independent
// LDS reads and the fma chain of cell_row_m, with the same operations in the same
// order (GMGSolver.cu:56-76).
struct CellPrep {
  int own;      // slot * 512 + cell, -1: no cell
  int nb[6];
  float cf[6];
  float dg, h, b;
  bool in;
};

__device__ __forceinline__ void cl_prep(const CoarseLDS& M, int s, int cc, CellPrep& P) {
  const CoarseSlot& S = M.S[s];
  const int i = cc >> 6, j = (cc >> 3) & 7, k = cc & 7;
  const uint8_t c0 = M.ct[s][cc];
  const bool inside[6] = {i > 0, i < 7, j > 0, j < 7, k > 0, k < 7};
  const int step[6] = {-64, 64, -8, 8, -1, 1};
  const int face[6] = {0, 3, 1, 4, 2, 5};
  P.own = s * 512 + cc;
  float d = 0.f;
#pragma unroll
  for (int e = 0; e < 6; e++) {
    const int q = inside[e] ? s : S.nbr[face[e]];
    const int nc = inside[e] ? cc + step[e] : cc - 7 * step[e];
    const uint8_t cn = q < 0 ? uint8_t(C_DIR) : M.ct[q < 0 ? 0 : q][nc];
    const float bc = inside[e] ? S.bc[6] : S.bc[face[e]];
    P.nb[e] = q < 0 ? CL_ZERO : q * 512 + nc;
    P.cf[e] = ((c0 | cn) & C_NEU) ? 0.f : bc;
    d = d + P.cf[e];
  }
  P.in = c0 & C_INT;
  P.h = S.h;
  P.dg = P.in ? ((d * P.h) * P.h) * P.h : 0.f;
  P.b = M.b[s][cc];
}

// A x of a prepared cell (cell_row_m's arithmetic); x0 returns the cell's value.
__device__ __forceinline__ float cl_ax(const CoarseLDS& M, const CellPrep& P, float& x0) {
  x0 = M.z[P.own];
  float xn[6];
#pragma unroll
  for (int e = 0; e < 6; e++) xn[e] = M.z[P.nb[e]];
  float sum = 0.f;
#pragma unroll
  for (int e = 0; e < 6; e++) sum = __builtin_fmaf(P.cf[e], x0 - xn[e], sum);
  return P.in ? ((sum * P.h) * P.h) * P.h : 0.f;
}

// The coarse levels 0..C of one V-cycle in one 1024-thread workgroup with every
// tile of those levels in LDS (the host keeps levels with GHOST tiles out of C
// and at most CL_MAX_PER_LEVEL tiles on a level): smoothing, restriction and
// prolongation per level and the bottom sweeps on level 0.  Each thread owns up
// to CL_CPT cells of each colour of the level being worked and keeps their
// prepared rows in registers.  One launch replaces ~10 launches per coarse
// level; the per-cell arithmetic is that of the tile kernels
// (GMGSolver.cu:216-376, bottom :432-433).
template <class V>
__global__ __launch_bounds__(CL_THREADS) void k_coarse_lds(Ctx c, const CoarsePlan* __restrict__ plans, VCycleArgs v,
                                                           V* Z, int flags) {
  __shared__ CoarseLDS M;
  __shared__ CoarsePlan cp;  // one workgroup per problem; in LDS, as its level table is indexed per level
  const int tid = threadIdx.x;
  if (tid == 0) cp = plans[blockIdx.x];
  __syncthreads();
  if ((flags & F_CHECK_DONE) && c.ps[cp.prob].done) return;
  if (tid < cp.nslots) M.S[tid] = cp.slots[tid];
  if (tid < 4) M.z[CL_ZERO + tid] = 0.f;
  __syncthreads();
  // cell types and rhs; z = 0
  for (int q = tid; q < cp.nslots * 128; q += CL_THREADS) {
    const int s = q >> 7, o = (q & 127) * 4;
    const size_t g = size_t(M.S[s].tile) * 512 + o;
    *reinterpret_cast<float4*>(&M.b[s][o]) = load_quad_split(vc_rhs<V>(c) + size_t(M.S[s].tile) * 512, o >> 3, o & 7);
    *reinterpret_cast<float4*>(&M.z[s * 512 + o]) = make_float4(0.f, 0.f, 0.f, 0.f);
    *reinterpret_cast<uint32_t*>(&M.ct[s][o]) = *reinterpret_cast<const uint32_t*>(c.ctype + g);
  }
  __syncthreads();
  CellPrep P[2][CL_CPT];  // [colour][m]
  // rows of this thread's cells of level l (cell m of colour col: q = tid + m * CL_THREADS)
  auto prep = [&](int l) {
    const int s0 = cp.lvl_begin[l], n = (cp.lvl_begin[l + 1] - s0) * 256;
#pragma unroll
    for (int col = 0; col < 2; col++)
#pragma unroll
      for (int m = 0; m < CL_CPT; m++) {
        const int q = tid + m * CL_THREADS;
        P[col][m].own = -1;
        if (q < n) {
          const int s = s0 + (q >> 8), r = q & 255;
          const int i = r >> 5, j = (r >> 2) & 7;
          cl_prep(M, s, (i << 6) | (j << 3) | (2 * (r & 3) + ((col + i + j) & 1)), P[col][m]);
        }
      }
  };
  // one red-black colour pass of the prepared level
  auto gs = [&](int col) {
#pragma unroll
    for (int m = 0; m < CL_CPT; m++) {
      const CellPrep& Q = P[col][m];
      if (Q.own >= 0) {
        float x0;
        const float ax = cl_ax(M, Q, x0);
        float delta = 0.f;
        if (Q.in && Q.dg != 0.f) delta = (Q.b - ax) / Q.dg;
        M.z[Q.own] = vround<V>(x0 + delta);
      }
    }
    __syncthreads();
  };
  // one damped-Jacobi sweep of the prepared level
  auto jacobi = [&]() {
    float out[2][CL_CPT];
#pragma unroll
    for (int col = 0; col < 2; col++)
#pragma unroll
      for (int m = 0; m < CL_CPT; m++) {
        const CellPrep& Q = P[col][m];
        out[col][m] = 0.f;
        if (Q.own >= 0) {
          float x0;
          const float ax = cl_ax(M, Q, x0);
          out[col][m] = x0;
          if (Q.in) out[col][m] = (Q.dg == 0.f) ? 0.f : x0 + (v.omega * (Q.b - ax)) / Q.dg;
        }
      }
    __syncthreads();
#pragma unroll
    for (int col = 0; col < 2; col++)
#pragma unroll
      for (int m = 0; m < CL_CPT; m++)
        if (P[col][m].own >= 0) M.z[P[col][m].own] = vround<V>(out[col][m]);
    __syncthreads();
  };
  auto smooth = [&](int order) {
    for (int it = 0; it < v.nu; it++) {
      if (v.smoother != 0) jacobi();
      else { gs(order); gs(1 - order); }
    }
  };
  for (int l = cp.C; l >= 1; l--) {
    prep(l);
    smooth(0);
    // residuals of level l, then their 2x2x2 sums into the rhs of level l-1
#pragma unroll
    for (int col = 0; col < 2; col++)
#pragma unroll
      for (int m = 0; m < CL_CPT; m++) {
        const CellPrep& Q = P[col][m];
        if (Q.own >= 0) {
          float x0;
          const float ax = cl_ax(M, Q, x0);
          (&M.t[0][0])[Q.own] = Q.b - ax;
        }
      }
    __syncthreads();
    const int s0 = cp.lvl_begin[l], ns = cp.lvl_begin[l + 1] - s0;
    for (int q = tid; q < ns * 64; q += CL_THREADS) {
      const int s = s0 + (q >> 6), lq = q & 63;
      const float sum = octet_sum(M.t[s], lq);
      const CoarseSlot& S = M.S[s];
      const int a = lq >> 4, bq = (lq >> 2) & 3, cq = lq & 3;
      const int ox = S.octant & 1, oy = (S.octant >> 1) & 1, oz = (S.octant >> 2) & 1;
      M.b[S.parent][((ox * 4 + a) << 6) | ((oy * 4 + bq) << 3) | (oz * 4 + cq)] = vround<V>(sum * v.ooa);
    }
    __syncthreads();
  }
  prep(0);
  const int it0 = v.coarsest / 2;
  for (int it = 0; it < it0; it++) { gs(0); gs(1); }
  for (int it = 0; it < it0; it++) { gs(1); gs(0); }
  for (int l = 1; l <= cp.C; l++) {
    const int s0 = cp.lvl_begin[l], ns = cp.lvl_begin[l + 1] - s0;
    for (int q = tid; q < ns * 512; q += CL_THREADS) {
      const int s = s0 + (q >> 9), cc = q & 511;
      const CoarseSlot& S = M.S[s];
      const int i = cc >> 6, j = (cc >> 3) & 7, k = cc & 7;
      const int ox = S.octant & 1, oy = (S.octant >> 1) & 1, oz = (S.octant >> 2) & 1;
      const float cz = M.z[S.parent * 512 + (((ox * 4 + (i >> 1)) << 6) | ((oy * 4 + (j >> 1)) << 3) | (oz * 4 + (k >> 1)))];
      M.z[s * 512 + cc] = (M.ct[s][cc] & C_INT) ? vround<V>(__builtin_fmaf(v.pcoef, cz, M.z[s * 512 + cc])) : 0.f;
    }
    __syncthreads();
    prep(l);
    smooth(1);
  }
  for (int q = tid; q < cp.nslots * 128; q += CL_THREADS) {
    const int s = q >> 7, o = (q & 127) * 4;
    store_quad_split(Z + size_t(M.S[s].tile) * 512, o >> 3, o & 7, *reinterpret_cast<const float4*>(&M.z[s * 512 + o]));
  }
}

// Empty kernel, for measuring the launch floor.
__global__ void k_empty(int* p) {
  if (p && threadIdx.x == 1024) *p = 0;
}

// ---------------------------------------------------------------------------
// CG kernels (rotated loop: V-cycle, gamma/beta, p, A p + alpha, x/r + check).
// Kernels write per-tile FP64 partials; k_finalize reduces them per problem.
// ---------------------------------------------------------------------------
// Per solve, for a candidate tile (LEAF, no ghost child, no INTERIOR cell,
// tr_inert_cand): it is inert when its initial residual is +0 in every cell.
// The residual then stays +0 (A p is 0 off INTERIOR cells), and so does every
// value the iterations would write there: the smoothers and the prolongation
// leave 0 on cells that are not INTERIOR, the residual and the restriction of
// zeros are 0, the operator writes 0, every dot partial is 0 (the slot holds
// the +0 written below) and x keeps its value (the updates are masked).  So
// its V-cycle vectors, A p and its octant block of the parent's V-cycle rhs
// (which only its restriction writes) are set to those values here, and the
// iteration kernels skip it (SKIP_INERT).
template <class V>
__device__ __forceinline__ void mark_inert(const Ctx& c, const TileRec& T, int lane, const float rv[8]) {
  bool zero = true;
  for (int k = 0; k < 8; k++) zero = zero && __float_as_uint(rv[k]) == 0u;
  const bool inert = __all(zero);
  if (lane == 0) c.inert[T.self] = inert ? 1 : 0;
  if (!inert) return;
  const float z8[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  store8(c.ap + size_t(T.self) * 512 + lane * 8, z8);
  store8_split(vc_z<V>(c) + size_t(T.self) * 512, lane, z8);
  store8_split(vc_zb<V>(c) + size_t(T.self) * 512, lane, z8);
  if (T.parent >= 0) vc_rhs<V>(c)[parent_cell_split(T, lane)] = from_f<V>(0.f);
}

// Init: x = 0, p = 0 on LEAF tiles, V-cycle rhs = b; partials of |b|^2
// (GMGSolver.cu:537-555).
template <class V>
__global__ __launch_bounds__(64) void k_cg_init(Ctx c, const TileRec* __restrict__ list, int n, int flags) {
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id();
  const size_t base = size_t(T.self) * 512 + lane * 8;
  float zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  store8(c.x + base, zero);
  store8(c.p + base, zero);
  float rv[8];
  load8(c.r + base, rv);
  store8_split(vc_rhs<V>(c) + size_t(T.self) * 512, lane, rv);
  double a = 0.0;
  for (int k = 0; k < 8; k++) a += double(rv[k]) * double(rv[k]);
  a = wave_sum(a);
  if (lane == 0) c.part[T.leaf_slot] = a;
  if (c.inert && tr_inert_cand(T)) mark_inert<V>(c, T, lane, rv);
}

// Warm-started init: r = b - A x0 on LEAF tiles (A x0 in AP, formed by the
// operator kernels), p = 0, V-cycle rhs = r; partials of |b|^2, against which
// the tolerance is measured.
template <class V>
__global__ __launch_bounds__(64) void k_cg_init_warm(Ctx c, const TileRec* __restrict__ list, int n, const float* __restrict__ B,
                                                     int flags) {
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id();
  const size_t base = size_t(T.self) * 512 + lane * 8;
  float bv[8], av[8], rv[8];
  load8(B + base, bv);
  load8(c.ap + base, av);
  for (int k = 0; k < 8; k++) rv[k] = bv[k] - av[k];
  float zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  store8(c.p + base, zero);
  store8(c.r + base, rv);
  store8_split(vc_rhs<V>(c) + size_t(T.self) * 512, lane, rv);
  double a = 0.0;
  for (int k = 0; k < 8; k++) a += double(bv[k]) * double(bv[k]);
  a = wave_sum(a);
  if (lane == 0) c.part[T.leaf_slot] = a;
  if (c.inert && tr_inert_cand(T)) mark_inert<V>(c, T, lane, rv);
}

// Pool build on the device: out[k] = tiles[ids[k]] (the worklists, records
// copied as 16-byte words), and the cell types permuted from the caller's tile
// order (src, tile t at t*512) to the canonical order (tile i from src tile
// ext[i]).
__global__ __launch_bounds__(256) void k_gather_tiles(const TileRec* __restrict__ tiles, const int32_t* __restrict__ ids,
                                                      int64_t n, TileRec* __restrict__ out) {
  constexpr int W = sizeof(TileRec) / 16;
  const int64_t words = n * W;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < words; k += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = k / W, w = k - r * W;
    reinterpret_cast<int4*>(out + r)[w] = reinterpret_cast<const int4*>(tiles + ids[r])[w];
  }
}
__global__ __launch_bounds__(64) void k_permute_ctype(const uint8_t* __restrict__ src, const int32_t* __restrict__ ext,
                                                      int n, uint8_t* __restrict__ dst) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const uint64_t* s8 = reinterpret_cast<const uint64_t*>(src + size_t(ext[i]) * 512);
  uint64_t* d8 = reinterpret_cast<uint64_t*>(dst + size_t(i) * 512);
  d8[threadIdx.x] = s8[threadIdx.x];
}
// A float channel between the caller's tile order and the canonical order:
// to_canonical: dst tile i = src tile ext[i]; otherwise dst tile ext[i] = src
// tile i.  One wavefront per tile, 32 bytes per lane.
__global__ __launch_bounds__(64) void k_permute_channel(const float* __restrict__ src, const int32_t* __restrict__ ext,
                                                        int n, float* __restrict__ dst, int to_canonical) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const size_t e = size_t(ext[i]) * 512, c = size_t(i) * 512;
  const float4* s4 = reinterpret_cast<const float4*>(src + (to_canonical ? e : c)) + 2 * threadIdx.x;
  float4* d4 = reinterpret_cast<float4*>(dst + (to_canonical ? c : e)) + 2 * threadIdx.x;
  const float4 a = s4[0], b = s4[1];
  d4[0] = a;
  d4[1] = b;
}

// Face masks of every cell of the listed tiles (Ctx::fmask): bit e of cell c is
// set when face e (x-, x+, y-, y+, z-, z+) is not cut, i.e. neither c nor its
// neighbour across e is NEUMANN (GMGSolver.cu:26-39; a neighbour outside the
// grid is DIRICHLET, GMGSolver.cu:120-122); bit 6 when c is INTERIOR.  Cell
// types do not change during a solve, so the masks are formed once.
__global__ __launch_bounds__(64) void k_face_mask(const uint8_t* __restrict__ ctype, uint8_t* __restrict__ fmask,
                                                  uint32_t* __restrict__ fsplit, const TileRec* __restrict__ list, int n) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  __shared__ uint64_t cs[100];
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t base = size_t(T.self) * 512 + lane * 8;
  const uint64_t own = load_ct8(ctype + base);
  const int pc = (i + 1) * 10 + (j + 1);
  cs[pc] = own;
  const uint64_t dir = 0x0101010101010101ull;
  if (lane < 32) {
    const int g = lane >> 3, t = lane & 7;
    const int nb = pick4(g, uni(T.nbr[0]), uni(T.nbr[3]), uni(T.nbr[1]), uni(T.nbr[4]));
    const int src = (g == 0) ? (56 + t) : (g == 1) ? t : (g == 2) ? (t * 8 + 7) : (t * 8);
    const int dst = (g == 0) ? (t + 1) : (g == 1) ? (91 + t) : (g == 2) ? ((t + 1) * 10) : ((t + 1) * 10 + 9);
    cs[dst] = nb < 0 ? dir : load_ct8(ctype + size_t(nb) * 512 + src * 8);
  }
  const uint8_t ckm = T.nbr[2] < 0 ? uint8_t(C_DIR) : ctype[size_t(T.nbr[2]) * 512 + lane * 8 + 7];
  const uint8_t ckp = T.nbr[5] < 0 ? uint8_t(C_DIR) : ctype[size_t(T.nbr[5]) * 512 + lane * 8];
  wave_sync();
  const uint64_t cim = cs[pc - 10], cip = cs[pc + 10], cjm = cs[pc - 1], cjp = cs[pc + 1];
  uint64_t out = 0;
  for (int k = 0; k < 8; k++) {
    const uint8_t c0 = ct_at(own, k);
    const uint8_t cn[6] = {ct_at(cim, k), ct_at(cip, k), ct_at(cjm, k), ct_at(cjp, k), k == 0 ? ckm : ct_at(own, k - 1),
                           k == 7 ? ckp : ct_at(own, k + 1)};
    uint32_t m = (c0 & C_INT) ? uint32_t(M_INT) : 0u;
    for (int e = 0; e < 6; e++)
      if (!((c0 | cn[e]) & C_NEU)) m |= 1u << e;
    out |= uint64_t(m) << (8 * k);
  }
  *reinterpret_cast<uint64_t*>(fmask + base) = out;
  const bool p0 = ((i + j) & 1) != 0;  // colour 0 holds the cells k = 2m + p0
  fsplit[size_t(T.self) * 128 + lane] = quad_mask(out, p0);
  fsplit[size_t(T.self) * 128 + 64 + lane] = quad_mask(out, !p0);
}

// Seeded update of a right-hand side, for a synthetic client of the serving
// loop: b = rho * b + sigma * u on INTERIOR cells of the listed (LEAF) tiles,
// u uniform on [-sqrt(3), sqrt(3)] (unit variance) from a counter hash of
// (seed, tile, cell).
__device__ __forceinline__ uint32_t hash32(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return uint32_t(x);
}
__global__ __launch_bounds__(64) void k_perturb_rhs(Ctx c, const TileRec* __restrict__ list, int n, float* __restrict__ B,
                                                    float rho, float sigma, uint64_t seed) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id();
  const size_t base = size_t(T.self) * 512 + lane * 8;
  float b[8];
  load8(B + base, b);
  const uint64_t ct = load_ct8(c.ctype + base);
  for (int k = 0; k < 8; k++) {
    if (!(ct_at(ct, k) & C_INT)) continue;
    const uint32_t h = hash32(seed * 0x9e3779b97f4a7c15ull + uint64_t(base + k));
    const float u = (float(h) * 2.3283064365386963e-10f * 2.0f - 1.0f) * 1.7320508075688772f;
    b[k] = rho * b[k] + sigma * u;
  }
  store8(B + base, b);
}

// Staged problem states to their slots: src[i] -> ps[map[i]], i < n (one
// workgroup per state).
__global__ __launch_bounds__(64) void k_ps_scatter(ProbState* __restrict__ ps, const ProbState* __restrict__ src,
                                                   const int32_t* __restrict__ map, int n) {
  static_assert(sizeof(ProbState) % 16 == 0, "ProbState is copied in 16-byte words");
  constexpr int W = sizeof(ProbState) / 16;
  const int i = blockIdx.x;
  if (i >= n) return;
  const int4* from = reinterpret_cast<const int4*>(src + i);
  int4* to = reinterpret_cast<int4*>(ps + map[i]);
  for (int k = threadIdx.x; k < W; k += blockDim.x) to[k] = from[k];
}

// Copy runs of tile records into an assembled worklist: run q copies
// counts[q] records from srcs[q] to dst + offs[q] (grid-stride over records).
__global__ __launch_bounds__(256) void k_gather_recs(const TileRec* const* __restrict__ srcs, const int32_t* __restrict__ offs,
                                                     const int32_t* __restrict__ counts, int nruns, TileRec* __restrict__ dst) {
  constexpr int W = sizeof(TileRec) / 16;
  for (int q = blockIdx.y; q < nruns; q += gridDim.y) {
    const int4* src = reinterpret_cast<const int4*>(srcs[q]);
    int4* out = reinterpret_cast<int4*>(dst + offs[q]);
    const int words = counts[q] * W;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < words; k += gridDim.x * blockDim.x) out[k] = src[k];
  }
}

// partials of gamma = r.z over LEAF tiles (GMGSolver.cu:634)
template <class V>
__global__ __launch_bounds__(64) void k_cg_dot_rz(Ctx c, const TileRec* __restrict__ list, int n, int flags) {
  TILE_KERNEL_PROLOGUE
  SKIP_INERT
  const int lane = lane_id();
  const size_t base = size_t(T.self) * 512 + lane * 8;
  float rv[8], zv[8];
  load8(c.r + base, rv);
  load8_split(vc_z<V>(c) + size_t(T.self) * 512, lane, zv);
  if (prob_done(c, T, flags)) return;
  double a = 0.0;
  for (int k = 0; k < 8; k++) a += double(rv[k]) * double(zv[k]);
  a = wave_sum(a);
  if (lane == 0) c.part[T.leaf_slot] = a;
}

// The previous iteration's x += alpha p (deferred from k_cg_update_xr, which
// then reads neither x nor p; the same fma on the same values, one iteration
// later, and k_cg_finish_x applies the last one), then p = z + beta p on
// INTERIOR cells of LEAF tiles (GMGSolver.cu:587-609 and 642-653).  The ghost
// children's p (the parent's INTERIOR values, PoissonGrid.cu:585-605) is not
// stored: the operator reads it from here (k_apply<true>).
template <class V>
__global__ __launch_bounds__(64) void k_cg_update_p(Ctx c, const TileRec* __restrict__ list, int n, int flags) {
  TILE_KERNEL_PROLOGUE
  SKIP_INERT
  const int lane = lane_id();
  const size_t base = size_t(T.self) * 512 + lane * 8;
  float pv[8], zv[8], xv[8];
  load8(c.p + base, pv);
  load8_split(vc_z<V>(c) + size_t(T.self) * 512, lane, zv);
  load8(c.x + base, xv);
  const uint64_t ct = tile_ct8(T, c.ctype + base);
  if (prob_done(c, T, flags)) return;
  const ProbState& ps = c.ps[T.prob];
  const float beta = float(ps.beta);
  const float alpha = float(ps.alpha);
  const bool upd_x = ps.iter > 0;  // an earlier iteration left its x update pending
  for (int k = 0; k < 8; k++)
    if (ct_at(ct, k) & C_INT) {
      if (upd_x) xv[k] = __builtin_fmaf(alpha, pv[k], xv[k]);
      pv[k] = __builtin_fmaf(beta, pv[k], zv[k]);
    }
  if (upd_x) store8(c.x + base, xv);
  store8(c.p + base, pv);
}

// A p on LEAF and GHOST tiles with the subtree view sl.  In CG mode (dot=1)
// LEAF tiles write A p and, when they have no ghost children, their p.Ap
// partial; GHOST tiles write the 2x2x2 sums of their rows for the parent
// (AccumulateToParents128 with all cell types, PoissonGrid.cu:664-710).  In
// channel mode (dot=0) every listed tile writes its rows to AX.
// PG (the CG iteration's operator on p): GHOST values from their parents
// (stage_issue_pg), as p is not copied into ghost tiles.
template <bool PG = false>
__global__ __launch_bounds__(64) void k_apply(Ctx c, const TileRec* __restrict__ list, int n, const float* __restrict__ X,
                                              float* __restrict__ AX, int sl, int diag, int dot, int flags) {
  TILE_KERNEL_PROLOGUE
  SKIP_INERT
  __shared__ TileLDS L;
  __shared__ float gl[512];
  Stencil S;
  float bcv[7];
  view_coeffs(T, sl, bcv);
  if (PG) {
    StagePG r = stage_issue_pg(c, T, X);
    stage_commit_pg(r, T, bcv, L, S);
  } else {
    stage_tile<false>(c, T, X, X, false, false, bcv, L, S);
  }
  if (prob_done(c, T, flags)) return;
  wave_sync();
  float ax[8], dg[8];
  if (diag) apply_pencil<false, true>(lds_nbrs(L), S, dg, ax);
  else apply_pencil<true, false>(lds_nbrs(L), S, ax, dg);
  const int lane = lane_id();
  const size_t base = size_t(T.self) * 512 + lane * 8;
  if (tr_type(T) == T_GHOST && dot) {
    const uint64_t ct = load_ct8(c.ctype + base);
    for (int k = 0; k < 8; k++) gl[lane * 8 + k] = (ct_at(ct, k) & (C_INT | C_DIR | C_NEU)) ? ax[k] : 0.f;
    wave_sync();
    c.gsum[size_t(T.gslot) * 64 + lane] = octet_sum(gl, lane);
    return;
  }
  store8(AX + base, ax);
  if (!dot || tr_type(T) != T_LEAF || tr_has_gchild(T)) return;
  double a = 0.0;
  for (int k = 0; k < 8; k++) a += double(S.own[k]) * double(ax[k]);
  a = wave_sum(a);
  if (lane == 0) c.part[T.leaf_slot] = a;
}

// LEAF tiles with ghost children: A p += ghost row sums (the += of
// AccumulateToParents128), then the p.Ap partial (GMGSolver.cu:580).
__global__ __launch_bounds__(64) void k_ghost_gather(Ctx c, const TileRec* __restrict__ list, int n, int flags) {
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t base = size_t(T.self) * 512 + lane * 8;
  float av[8], pv[8];
  load8(c.ap + base, av);
  load8(c.p + base, pv);
  float4 e[2];
  int gs[2];
  for (int h = 0; h < 2; h++) {
    gs[h] = list[w].gcs[(i >> 2) | ((j >> 2) << 1) | (h << 2)];  // lane-dependent index: read from memory
    if (gs[h] >= 0) e[h] = *reinterpret_cast<const float4*>(c.gsum + size_t(gs[h]) * 64 + (i & 3) * 16 + (j & 3) * 4);
  }
  if (prob_done(c, T, flags)) return;
  for (int h = 0; h < 2; h++) {
    if (gs[h] < 0) continue;
    av[4 * h + 0] = av[4 * h + 0] + e[h].x;
    av[4 * h + 1] = av[4 * h + 1] + e[h].y;
    av[4 * h + 2] = av[4 * h + 2] + e[h].z;
    av[4 * h + 3] = av[4 * h + 3] + e[h].w;
  }
  store8(c.ap + base, av);
  double a = 0.0;
  for (int k = 0; k < 8; k++) a += double(pv[k]) * double(av[k]);
  a = wave_sum(a);
  if (lane == 0) c.part[T.leaf_slot] = a;
}

// r -= alpha A p on INTERIOR cells of LEAF tiles (GMGSolver.cu:587-609); r is
// also the next V-cycle's rhs; partials of |r|^2.  The iteration's x += alpha p
// is left to the next k_cg_update_p (or k_cg_finish_x).
template <class V>
__global__ __launch_bounds__(64) void k_cg_update_xr(Ctx c, const TileRec* __restrict__ list, int n, int flags) {
  TILE_KERNEL_PROLOGUE
  SKIP_INERT
  const int lane = lane_id();
  const size_t base = size_t(T.self) * 512 + lane * 8;
  float rv[8], av[8];
  load8(c.r + base, rv);
  load8(c.ap + base, av);
  const uint64_t ct = tile_ct8(T, c.ctype + base);
  if (prob_done(c, T, flags)) return;
  const float alpha = float(c.ps[T.prob].alpha);
  for (int k = 0; k < 8; k++)
    if (ct_at(ct, k) & C_INT) rv[k] = __builtin_fmaf(-alpha, av[k], rv[k]);
  store8(c.r + base, rv);
  store8_split(vc_rhs<V>(c) + size_t(T.self) * 512, lane, rv);
  double a = 0.0;
  for (int k = 0; k < 8; k++) a += double(rv[k]) * double(rv[k]);
  a = wave_sum(a);
  if (lane == 0) c.part[T.leaf_slot] = a;
}

// The last iteration's pending x += alpha p on INTERIOR cells of LEAF tiles,
// once after a problem's iterations end (every problem of the list, done or
// not; none when no iteration ran).
__global__ __launch_bounds__(64) void k_cg_finish_x(Ctx c, const TileRec* __restrict__ list, int n, int flags) {
  TILE_KERNEL_PROLOGUE
  SKIP_INERT
  const int lane = lane_id();
  const size_t base = size_t(T.self) * 512 + lane * 8;
  float pv[8], xv[8];
  load8(c.p + base, pv);
  load8(c.x + base, xv);
  const uint64_t ct = tile_ct8(T, c.ctype + base);
  const ProbState& ps = c.ps[T.prob];
  if (ps.iter <= 0) return;
  const float alpha = float(ps.alpha);
  for (int k = 0; k < 8; k++)
    if (ct_at(ct, k) & C_INT) xv[k] = __builtin_fmaf(alpha, pv[k], xv[k]);
  store8(c.x + base, xv);
}

// Sum of a problem's per-tile partials in the fixed order (lane q adds
// partials q, q+64, ... in sequence, then the butterfly), staged through LDS by
// the whole workgroup.  Each thread issues all its loads of a chunk before the
// first wait (addresses clamped into the range, so every load is in bounds),
// and the summing wave reads its LDS values in groups of FIN_GROUP before
// adding them in sequence: a chunk costs one memory round trip, not one per
// element.
constexpr int FIN_THREADS = 256;
constexpr int FIN_CHUNK = 4096;
constexpr int FIN_PER_THREAD = FIN_CHUNK / FIN_THREADS;
constexpr int FIN_GROUP = 16;
__device__ __forceinline__ double block_sum_partials(const double* __restrict__ part, int begin, int count, double* lds) {
  const int tid = threadIdx.x;
  double a = 0.0;
  for (int c0 = 0; c0 < count; c0 += FIN_CHUNK) {
    const int m = min(FIN_CHUNK, count - c0);  // >= 1
    const double* src = part + begin + c0;
    double v[FIN_PER_THREAD];
#pragma unroll
    for (int u = 0; u < FIN_PER_THREAD; u++) v[u] = src[min(tid + u * FIN_THREADS, m - 1)];
    __syncthreads();  // the previous chunk's sums are done with lds
#pragma unroll
    for (int u = 0; u < FIN_PER_THREAD; u++)
      if (tid + u * FIN_THREADS < m) lds[tid + u * FIN_THREADS] = v[u];
    __syncthreads();
    if (tid < 64) {
      // lane tid continues its sequence tid, tid+64, ... across chunks (FIN_CHUNK is a multiple of 64)
      const int len = (m - tid + 63) >> 6;  // values of this lane in the chunk
      for (int g0 = 0; g0 < len; g0 += FIN_GROUP) {
        double w[FIN_GROUP];
#pragma unroll
        for (int u = 0; u < FIN_GROUP; u++) w[u] = lds[min(tid + (g0 + u) * 64, FIN_CHUNK - 1)];
#pragma unroll
        for (int u = 0; u < FIN_GROUP; u++)
          if (g0 + u < len) a += w[u];
      }
    }
  }
  if (tid < 64) a = wave_sum(a);
  return a;
}

// Per-problem finalisation of a reduction over the LEAF-tile partials.
enum : int { FIN_BB = 0, FIN_RZ = 1, FIN_PAP = 2, FIN_RR = 3, FIN_MEAN = 4, FIN_IRBB = 5, FIN_SCALE = 6 };
// range_count >= 0: the launch's one problem has its partials at range_begin ..
// range_begin + range_count - 1 (its leaf range), given by the host so that
// their loads go out with the problem state's rather than after it.
__global__ __launch_bounds__(FIN_THREADS) void k_finalize(Ctx c, int mode, int flags, int range_begin = 0,
                                                          int range_count = -1) {
  __shared__ double lds[FIN_CHUNK];
  ProbState* ps = c.ps + (c.pmap ? c.pmap[blockIdx.x] : int(blockIdx.x));
  const bool given = range_count >= 0;
  if (!given) {
    if (mode != FIN_BB && (flags & F_CHECK_DONE) && ps->done) return;
    if (mode == FIN_MEAN && !ps->pure_neumann) return;
  }
  const int begin = given ? range_begin : ps->leaf_begin;
  const int count = given ? range_count : ps->leaf_count;
  const double s = block_sum_partials(c.part, begin, count, lds);
  if (given) {  // the same tests, after the sum (whose value is then dropped)
    if (mode != FIN_BB && (flags & F_CHECK_DONE) && ps->done) return;
    if (mode == FIN_MEAN && !ps->pure_neumann) return;
  }
  double q = 0.0;
  if (mode == FIN_MEAN) q = block_sum_partials(c.part2, begin, count, lds);
  if (threadIdx.x != 0) return;
  switch (mode) {
    case FIN_BB: {  // GMGSolver.cu:542-554
      ps->bb = s;
      ps->rr = s;
      double thr = ps->tol2 * s;
      thr = thr > ps->thr_abs ? thr : ps->thr_abs;
      ps->thr = thr > 2.2250738585072014e-308 ? thr : 2.2250738585072014e-308;
      ps->iter = 0;
      ps->converged = (s == 0.0) ? 1 : 0;
      ps->done = (s == 0.0 || ps->max_iters <= 0) ? 1 : 0;
      ps->gamma = 0.0;
      c.dflag[blockIdx.x] = ps->done;
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
    case FIN_RR: {  // GMGSolver.cu:611-619 (checked every sync_stride iterations)
      ps->rr = s;
      int it = ps->iter + 1;
      ps->iter = it;
      const int cs = ps->check_stride;
      if (cs > 0 && it % cs == 0 && s < ps->thr) { ps->converged = 1; ps->done = 1; }
      else if (it >= ps->max_iters) ps->done = 1;
      c.dflag[blockIdx.x] = ps->done;
      break;
    }
    case FIN_MEAN:
      ps->mean = s / q;
      break;
    case FIN_IRBB:  // |b|^2 of an FP64 refinement's right-hand side and the inner solves' threshold floor
      ps->bb_b = s;
      ps->thr_abs = ps->tol_abs2 * s;
      break;
    case FIN_SCALE: {  // float16 storage: the right-hand side's scale 2^k, from |b|^2 / cells = 2^e f,
                       // f in [1, 2): k = -floor(e/2) - 4 (root mean square near 2^-4)
      int k = 0;
      if (s > 0.0) {
        const int e = ilogb(s / (512.0 * double(ps->leaf_count)));
        const int half = (e >= 0) ? e / 2 : -((-e + 1) / 2);
        k = min(120, max(-120, -half - 4));
      }
      ps->scale_exp = k;
      ps->thr_abs = ldexp(ps->thr_abs, 2 * k);
      break;
    }
  }
}

// Pure-Neumann re-centring (MeanAsync + ReCenterLeafVoxels,
// PoissonGrid.cu:231-325, PoissonSolver.cu:356-368): mean of r over all cells
// of LEAF tiles divided by the INTERIOR count, subtracted on INTERIOR cells.
__global__ __launch_bounds__(64) void k_mean(Ctx c, const TileRec* __restrict__ list, int n, const float* __restrict__ X, int flags) {
  TILE_KERNEL_PROLOGUE
  if (prob_done(c, T, flags)) return;
  ProbState* ps = c.ps + T.prob;
  if (!ps->pure_neumann) return;
  const int lane = lane_id();
  const size_t base = size_t(T.self) * 512 + lane * 8;
  float v[8];
  load8(X + base, v);
  const uint64_t ct = load_ct8(c.ctype + base);
  double a = 0.0, q = 0.0;
  for (int k = 0; k < 8; k++) { a += double(v[k]); q += (ct_at(ct, k) & C_INT) ? 1.0 : 0.0; }
  a = wave_sum(a);
  q = wave_sum(q);
  if (lane == 0) { c.part[T.leaf_slot] = a; c.part2[T.leaf_slot] = q; }
}
template <class V>
__global__ __launch_bounds__(64) void k_subtract_mean(Ctx c, const TileRec* __restrict__ list, int n, float* __restrict__ X,
                                                      int write_rhsv, int flags) {
  TILE_KERNEL_PROLOGUE
  if (prob_done(c, T, flags)) return;
  ProbState* ps = c.ps + T.prob;
  if (!ps->pure_neumann) return;
  const double mean = ps->mean;
  const size_t base = size_t(T.self) * 512 + lane_id() * 8;
  float v[8];
  load8(X + base, v);
  const uint64_t ct = load_ct8(c.ctype + base);
  for (int k = 0; k < 8; k++) if (ct_at(ct, k) & C_INT) v[k] = float(double(v[k]) - mean);
  store8(X + base, v);
  if (write_rhsv) store8_split(vc_rhs<V>(c) + size_t(T.self) * 512, lane_id(), v);
}

// ---------------------------------------------------------------------------
// Channel-level kernels for the operator API
// ---------------------------------------------------------------------------
// Ghost x from the parent's INTERIOR value (PoissonGrid.cu:585-605).
__global__ __launch_bounds__(64) void k_propagate_ghosts(Ctx c, const TileRec* __restrict__ list, int n,
                                                         const float* __restrict__ XC, float* __restrict__ XF) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t base = size_t(T.self) * 512 + lane * 8;
  const int oct = tr_octant(T), ox = oct & 1, oy = (oct >> 1) & 1, oz = (oct >> 2) & 1;
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
__global__ __launch_bounds__(64) void k_accumulate_ghosts(Ctx c, const TileRec* __restrict__ list, int n, float* __restrict__ AX) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id();
  const size_t base = size_t(T.self) * 512 + lane * 8;
  __shared__ float gl[512];
  float v[8];
  load8(AX + base, v);
  const uint64_t ct = load_ct8(c.ctype + base);
  for (int k = 0; k < 8; k++) gl[lane * 8 + k] = (ct_at(ct, k) & (C_INT | C_DIR | C_NEU)) ? v[k] : 0.f;
  wave_sync();
  float s = octet_sum(gl, lane);
  size_t pcell = parent_cell(T, lane);
  AX[pcell] = AX[pcell] + s * 1.0f;
}
// Per-tile dot partials over all cells of the listed tiles, then one total.
// FP64 sum of a[c] * b[c] over the 512 cells of one tile: each lane over its
// pencil in cell order, then the xor butterfly across the wavefront.
__device__ __forceinline__ double tile_dot(const float* __restrict__ A, const float* __restrict__ Bv, int tile) {
  const size_t base = size_t(tile) * 512 + lane_id() * 8;
  float a[8], b[8];
  load8(A + base, a);
  load8(Bv + base, b);
  double s = 0.0;
  for (int k = 0; k < 8; k++) s += double(a[k]) * double(b[k]);
  return wave_sum(s);
}
__global__ __launch_bounds__(64) void k_dot_tiles(Ctx c, const TileRec* __restrict__ list, int n, const float* __restrict__ A,
                                                  const float* __restrict__ Bv, double* __restrict__ part) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  const double s = tile_dot(A, Bv, T.self);
  if (lane_id() == 0) part[w] = s;
}
// The indices i < n (ascending) of the tiles whose type is in `types`, written
// to sel[0 ..); one workgroup of SEL_THREADS walks the records in chunks, each
// chunk's ranks from wavefront ballots and a scan of the wavefront counts.
constexpr int SEL_THREADS = 1024;
__global__ __launch_bounds__(SEL_THREADS) void k_select_tiles(const TileRec* __restrict__ tiles, int n, int types,
                                                              int32_t* __restrict__ sel) {
  __shared__ int wcount[SEL_THREADS / 64];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  int base = 0;
  for (int c0 = 0; c0 < n; c0 += SEL_THREADS) {
    const int i = c0 + tid;
    const bool take = i < n && (tr_type(tiles[i]) & types) != 0;
    const uint64_t bal = __ballot(take);
    const int below = __popcll(bal & ((uint64_t(1) << lane) - 1));
    if (lane == 0) wcount[wave] = __popcll(bal);
    __syncthreads();
    int before = 0, total = 0;
    for (int v = 0; v < SEL_THREADS / 64; v++) {
      before += v < wave ? wcount[v] : 0;
      total += wcount[v];
    }
    if (take) sel[base + before + below] = i;
    base += total;
    __syncthreads();  // wcount is rewritten by the next chunk
  }
}
// Tile partials of the selected tiles: partial q of tile sel[q] (q < count).
__global__ __launch_bounds__(64) void k_dot_selected(const int32_t* __restrict__ sel, int count, const float* __restrict__ A,
                                                     const float* __restrict__ Bv, double* __restrict__ part) {
  const int q = blockIdx.x;
  if (q >= count) return;
  const double s = tile_dot(A, Bv, sel[q]);
  if (lane_id() == 0) part[q] = s;
}
__global__ __launch_bounds__(FIN_THREADS) void k_sum_partials(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double lds[FIN_CHUNK];
  double s = block_sum_partials(part, 0, n, lds);
  if (threadIdx.x == 0) *out = s;
}
// Copy whole tiles of a list from one channel to another.
__global__ __launch_bounds__(64) void k_copy_tiles(const TileRec* __restrict__ list, int n, const float* __restrict__ src,
                                                   float* __restrict__ dst) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  const size_t base = size_t(T.self) * 512 + lane_id() * 8;
  float v[8];
  load8(src + base, v);
  store8(dst + base, v);
}
// Natural <-> colour-split order of whole tiles (channels handed to the
// operator API are natural; the V-cycle kernels work on split arrays).
__global__ __launch_bounds__(64) void k_to_split(const TileRec* __restrict__ list, int n, const float* __restrict__ src,
                                                 float* __restrict__ dst) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  float v[8];
  load8(src + size_t(T.self) * 512 + lane_id() * 8, v);
  store8_split(dst + size_t(T.self) * 512, lane_id(), v);
}
__global__ __launch_bounds__(64) void k_from_split(const TileRec* __restrict__ list, int n, const float* __restrict__ src,
                                                   float* __restrict__ dst) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  float v[8];
  load8_split(src + size_t(T.self) * 512, lane_id(), v);
  store8(dst + size_t(T.self) * 512 + lane_id() * 8, v);
}
// ---------------------------------------------------------------------------
// AMGSolver (AMGSolver.cu): stored coefficients and their V-cycle, natural
// cell order, FP32.  Coefficient arrays: off[a] = the cell's minus-face entry
// along axis a, diag; cell types of the AMG hierarchy in amgct (NONLEAF cells
// retyped by the coarsening, the others as given).
// ---------------------------------------------------------------------------
struct AmgC {
  float* off[3];
  float* diag;
  uint8_t* ct;
};
// AMGSolver.cu:17-36
__device__ __forceinline__ float amg_coeff(float h, uint8_t t0, uint8_t t1, uint8_t c0, uint8_t c1) {
  const int both = (t0 & T_LEAF) && (t1 & T_LEAF);
  const int lg = ((t0 & T_LEAF) && (t1 & T_GHOST)) || ((t0 & T_GHOST) && (t1 & T_LEAF));
  const float odh = (1.f / h) * float(both) + (1.f / (1.5f * h)) * float(lg);
  const int hn = (c0 & C_NEU) || (c1 & C_NEU);
  return hn ? 0.f : odh / h;
}
// face neighbour tile f (x-,y-,z-,x+,y+,z+) of T: index (-1: none) and type
__device__ __forceinline__ int amg_nbr(const TileRec& T, int f) { return uni(T.nbr[f]); }

// Step 1 of CoarsenTypesAndAMGCoeffs (AMGSolver.cu:118-131): a GHOST cell takes
// its parent cell's type (a GHOST tile's parent always exists).
__global__ __launch_bounds__(64) void k_amg_ghost_types(const TileRec* __restrict__ list, int n,
                                                        const uint8_t* __restrict__ ctype, uint8_t* __restrict__ amgct) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const int o = tr_octant(T);
  const size_t pb = size_t(T.parent) * 512 + (((o & 1) * 4 + (i >> 1)) * 8 + (((o >> 1) & 1) * 4 + (j >> 1))) * 8 +
                    ((o >> 2) & 1) * 4;
  for (int k = 0; k < 8; k++) amgct[size_t(T.self) * 512 + lane * 8 + k] = ctype[pb + (k >> 1)];
}

// CalculateAMGCoefficients (AMGSolver.cu:40-107) on LEAF|GHOST tiles: lane
// (i,j) owns cells k = 0..7; neighbour cell types from the own tile or the
// face neighbour (none: DIRICHLET cell, the tile's own type).
__global__ __launch_bounds__(64) void k_amg_coeffs(const TileRec* __restrict__ list, int n, AmgC A) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t tb = size_t(T.self) * 512;
  const uint8_t t0 = tr_type(T);
  const float h = T.h;
  const uint8_t* ctype = A.ct;
  const uint64_t own = load_ct8(ctype + tb + lane * 8);
  // the four in-plane neighbour pencils' cell types and tile types
  uint64_t nct[4];
  uint8_t ntt[4];
  const int pi[4] = {i - 1, i + 1, i, i};
  const int pj[4] = {j, j, j - 1, j + 1};
  const int pf[4] = {0, 3, 1, 4};
  for (int q = 0; q < 4; q++) {
    const bool inside = pi[q] >= 0 && pi[q] < 8 && pj[q] >= 0 && pj[q] < 8;
    if (inside) { nct[q] = load_ct8(ctype + tb + (pi[q] * 8 + pj[q]) * 8); ntt[q] = t0; continue; }
    const int nb = amg_nbr(T, pf[q]);
    if (nb < 0) { nct[q] = 0x0101010101010101ull; ntt[q] = t0; continue; }  // DIRICHLET, own type
    nct[q] = load_ct8(ctype + size_t(nb) * 512 + (((pi[q] + 8) & 7) * 8 + ((pj[q] + 8) & 7)) * 8);
    ntt[q] = tr_ntype(T, pf[q]);
  }
  uint8_t kmc = C_DIR, kpc = C_DIR, kmt = t0, kpt = t0;
  {
    const int nb = amg_nbr(T, 2);
    if (nb >= 0) { kmc = ctype[size_t(nb) * 512 + lane * 8 + 7]; kmt = tr_ntype(T, 2); }
    const int np = amg_nbr(T, 5);
    if (np >= 0) { kpc = ctype[size_t(np) * 512 + lane * 8]; kpt = tr_ntype(T, 5); }
  }
  float out[4][8];
  for (int k = 0; k < 8; k++) {
    const uint8_t c0 = ct_at(own, k);
    const uint8_t c1s[6] = {ct_at(nct[0], k), ct_at(nct[1], k), ct_at(nct[2], k), ct_at(nct[3], k),
                            k > 0 ? ct_at(own, k - 1) : kmc, k < 7 ? ct_at(own, k + 1) : kpc};
    const uint8_t t1s[6] = {ntt[0], ntt[1], ntt[2], ntt[3], k > 0 ? t0 : kmt, k < 7 ? t0 : kpt};
    float diag = 0.f, off[3] = {0.f, 0.f, 0.f};
    for (int e = 0; e < 6; e++) {  // x-, x+, y-, y+, z-, z+
      const float coeff = ((amg_coeff(h, t0, t1s[e], c0, c1s[e]) * h) * h) * h;
      if ((e & 1) == 0 && c1s[e] == C_INT) off[e >> 1] = -coeff;
      diag = diag + coeff;
    }
    const bool in = c0 == C_INT;
    for (int a = 0; a < 3; a++) out[a][k] = in ? off[a] : 0.f;
    out[3][k] = in ? diag : 0.f;
  }
  for (int a = 0; a < 3; a++) store8(A.off[a] + tb + lane * 8, out[a]);
  store8(A.diag + tb + lane * 8, out[3]);
}

// Step 3 of CoarsenTypesAndAMGCoeffs (AMGSolver.cu:137-207) from the child
// side: the tiles of one level (children) each fill the 4x4x4 cells of their
// parent's octant, lane = parent cell.  LEAF|NONLEAF children: the NONLEAF
// parent's cell type and the Galerkin sums times one_over_alpha; GHOST
// children: the sums added to the LEAF parent's own coefficients.  Sums in the
// reference's order (children ci, cj, ck lexicographic).
__global__ __launch_bounds__(64) void k_amg_coarsen(const TileRec* __restrict__ list, int n, AmgC A, float alpha) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id();
  const int qi = lane >> 4, qj = (lane >> 2) & 3, qk = lane & 3;
  const size_t tb = size_t(T.self) * 512;
  float cod[3][8];
  float dsum = 0.f;
  int interior = 0;
  for (int ci = 0; ci < 2; ci++)
    for (int cj = 0; cj < 2; cj++)
      for (int ck = 0; ck < 2; ck++) {
        const int q = (ci * 2 + cj) * 2 + ck;
        const size_t cc = tb + ((2 * qi + ci) * 8 + (2 * qj + cj)) * 8 + (2 * qk + ck);
        dsum = dsum + A.diag[cc];
        interior += A.ct[cc] == C_INT;
        for (int a = 0; a < 3; a++) cod[a][q] = A.off[a][cc];
      }
  float offs[3] = {0.f, 0.f, 0.f};
  for (int ci = 0; ci < 2; ci++)
    for (int cj = 0; cj < 2; cj++)
      for (int ck = 0; ck < 2; ck++) {
        const int q = (ci * 2 + cj) * 2 + ck;
        if (ci == 0) offs[0] = offs[0] + cod[0][q];
        if (ci == 1) dsum = dsum + 2.f * cod[0][q];
        if (cj == 0) offs[1] = offs[1] + cod[1][q];
        if (cj == 1) dsum = dsum + 2.f * cod[1][q];
        if (ck == 0) offs[2] = offs[2] + cod[2][q];
        if (ck == 1) dsum = dsum + 2.f * cod[2][q];
      }
  const int o = tr_octant(T);
  const size_t pc = size_t(T.parent) * 512 + (((o & 1) * 4 + qi) * 8 + (((o >> 1) & 1) * 4 + qj)) * 8 + (((o >> 2) & 1) * 4 + qk);
  if (tr_type(T) == T_GHOST) {
    for (int a = 0; a < 3; a++) A.off[a][pc] = A.off[a][pc] + offs[a];
    A.diag[pc] = A.diag[pc] + dsum;
  } else {
    A.ct[pc] = interior > 0 ? C_INT : C_DIR;
    for (int a = 0; a < 3; a++) A.off[a][pc] = offs[a] * alpha;
    A.diag[pc] = dsum * alpha;
  }
}

// VCycleAMG's first step (AMGSolver.cu:616-628): x = 0, rhs = f (0 on GHOST tiles).
__global__ __launch_bounds__(64) void k_amg_init(const TileRec* __restrict__ list, int n, float* __restrict__ X,
                                                 const float* __restrict__ F, float* __restrict__ RHS, int flags) {
  TILE_KERNEL_PROLOGUE
  const size_t b = size_t(T.self) * 512 + lane_id() * 8;
  float v[8], z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (tr_type(T) == T_GHOST) store8(RHS + b, z);
  else { load8(F + b, v); store8(RHS + b, v); }
  store8(X + b, z);
}

// A x with the stored coefficients for the lane's 8 cells (negativeLapSameLevel,
// AMGSolver.cu:232-250): x0 d + sum over x-, x+, y-, y+, z-, z+ of the face's
// entry times the neighbour's x, the plus faces' entries read from the plus
// neighbour's minus-face entry (0 where there is no neighbour tile).
__device__ __forceinline__ void amg_rows(const TileRec& T, const float* __restrict__ X, const AmgC& A, float x0[8],
                                         float ax[8]) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t tb = size_t(T.self) * 512;
  load8(X + tb + lane * 8, x0);
  float xs[4][8], os[2][8];  // x of the x-, x+, y-, y+ pencils; off_x of x+, off_y of y+
  const int pi[4] = {i - 1, i + 1, i, i};
  const int pj[4] = {j, j, j - 1, j + 1};
  const int pf[4] = {0, 3, 1, 4};
  for (int q = 0; q < 4; q++) {
    const bool inside = pi[q] >= 0 && pi[q] < 8 && pj[q] >= 0 && pj[q] < 8;
    const int nb = inside ? int(T.self) : amg_nbr(T, pf[q]);
    const size_t pb = size_t(nb) * 512 + (((pi[q] + 8) & 7) * 8 + ((pj[q] + 8) & 7)) * 8;
    if (nb < 0) {
      for (int k = 0; k < 8; k++) xs[q][k] = 0.f;
      if (q == 1) for (int k = 0; k < 8; k++) os[0][k] = 0.f;
      if (q == 3) for (int k = 0; k < 8; k++) os[1][k] = 0.f;
      continue;
    }
    load8(X + pb, xs[q]);
    if (q == 1) load8(A.off[0] + pb, os[0]);
    if (q == 3) load8(A.off[1] + pb, os[1]);
  }
  float ox[8], oy[8], oz[8], dg[8];
  load8(A.off[0] + tb + lane * 8, ox);
  load8(A.off[1] + tb + lane * 8, oy);
  load8(A.off[2] + tb + lane * 8, oz);
  load8(A.diag + tb + lane * 8, dg);
  float xkm = 0.f, xkp = 0.f, ozp = 0.f;
  {
    const int nb = amg_nbr(T, 2);
    if (nb >= 0) xkm = X[size_t(nb) * 512 + lane * 8 + 7];
    const int np = amg_nbr(T, 5);
    if (np >= 0) { xkp = X[size_t(np) * 512 + lane * 8]; ozp = A.off[2][size_t(np) * 512 + lane * 8]; }
  }
  for (int k = 0; k < 8; k++) {
    float sum = x0[k] * dg[k];
    sum = __builtin_fmaf(ox[k], xs[0][k], sum);
    sum = __builtin_fmaf(os[0][k], xs[1][k], sum);
    sum = __builtin_fmaf(oy[k], xs[2][k], sum);
    sum = __builtin_fmaf(os[1][k], xs[3][k], sum);
    sum = __builtin_fmaf(oz[k], k > 0 ? x0[k - 1] : xkm, sum);
    sum = __builtin_fmaf(k < 7 ? oz[k + 1] : ozp, k < 7 ? x0[k + 1] : xkp, sum);
    ax[k] = sum;
  }
}

// One colour of GaussSeidelAMG (AMGSolver.cu:540-578) on LEAF|NONLEAF tiles
// (D = 0 leaves the cell unchanged, as in the GMG smoother).
__global__ __launch_bounds__(64) void k_amg_gs(Ctx c, const TileRec* __restrict__ list, int n, float* __restrict__ X,
                                               const float* __restrict__ RHS, AmgC A, int color, int flags) {
  TILE_KERNEL_PROLOGUE
  float x0[8], ax[8];
  amg_rows(T, X, A, x0, ax);
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t b = size_t(T.self) * 512 + lane * 8;
  float rhs[8], dg[8];
  load8(RHS + b, rhs);
  load8(A.diag + b, dg);
  const uint64_t ct = load_ct8(A.ct + b);
  if (prob_done(c, T, flags)) return;
  wave_sync();  // every lane has read its neighbours before any lane writes
  float out[8];
  for (int k = 0; k < 8; k++) {
    out[k] = x0[k];
    if (((i + j + k) & 1) != color) continue;
    const float delta = (ct_at(ct, k) & C_INT) ? ((dg[k] != 0.f) ? (rhs[k] - ax[k]) / dg[k] : 0.f) : 0.f;
    out[k] = x0[k] + delta;
  }
  store8(X + b, out);
}

// The residual of VCycleAMG (AMGSolver.cu:638-642): tmp = rhs - A x on
// INTERIOR cells, 0 elsewhere, on every tile of a level.
__global__ __launch_bounds__(64) void k_amg_residual(Ctx c, const TileRec* __restrict__ list, int n,
                                                     const float* __restrict__ X, const float* __restrict__ RHS,
                                                     float* __restrict__ TMP, AmgC A, int flags) {
  TILE_KERNEL_PROLOGUE
  float x0[8], ax[8], rhs[8];
  amg_rows(T, X, A, x0, ax);
  const size_t b = size_t(T.self) * 512 + lane_id() * 8;
  load8(RHS + b, rhs);
  const uint64_t ct = load_ct8(A.ct + b);
  if (prob_done(c, T, flags)) return;
  float out[8];
  for (int k = 0; k < 8; k++) out[k] = (ct_at(ct, k) == C_INT) ? rhs[k] - ax[k] : 0.f;
  store8(TMP + b, out);
}

// Restrict (PoissonSolver.cu:292-316) from the child side: the LEAF|NONLEAF
// tiles of a level fill their NONLEAF parent's octant, rhs = sum over the 8
// children (x fastest) of their INTERIOR residuals, times one_over_alpha.
__global__ __launch_bounds__(64) void k_amg_restrict(Ctx c, const TileRec* __restrict__ list, int n,
                                                     const float* __restrict__ TMP, float* __restrict__ RHS, AmgC A,
                                                     float alpha, int flags) {
  TILE_KERNEL_PROLOGUE
  if (prob_done(c, T, flags)) return;
  const int lane = lane_id();
  const int qi = lane >> 4, qj = (lane >> 2) & 3, qk = lane & 3;
  const size_t tb = size_t(T.self) * 512;
  float sum = 0.f;
  for (int s = 0; s < 8; s++) {
    const size_t cc = tb + ((2 * qi + (s & 1)) * 8 + (2 * qj + ((s >> 1) & 1))) * 8 + (2 * qk + (s >> 2));
    sum = sum + ((A.ct[cc] & C_INT) ? TMP[cc] : 0.f);
  }
  const int o = tr_octant(T);
  RHS[size_t(T.parent) * 512 + (((o & 1) * 4 + qi) * 8 + (((o >> 1) & 1) * 4 + qj)) * 8 + (((o >> 2) & 1) * 4 + qk)] =
      sum * alpha;
}

// VCycleAMG's rhs on the LEAF and GHOST tiles of the coarser level
// (AMGSolver.cu:652-656): LEAF rhs = f plus the 8-cell sums of the GHOST
// children's residuals (AccumulateToParents, PoissonGrid.cu:721-755), GHOST rhs = 0.
__global__ __launch_bounds__(64) void k_amg_leaf_rhs(Ctx c, const TileRec* __restrict__ list, int n,
                                                     const float* __restrict__ TMP, const float* __restrict__ F,
                                                     float* __restrict__ RHS, int flags) {
  TILE_KERNEL_PROLOGUE
  const uint8_t ty = tr_type(T);
  if (ty == T_NONLEAF || prob_done(c, T, flags)) return;
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t b = size_t(T.self) * 512 + lane * 8;
  float out[8];
  if (ty == T_GHOST) {
    for (int k = 0; k < 8; k++) out[k] = 0.f;
  } else {
    float f[8];
    load8(F + b, f);
    for (int k = 0; k < 8; k++) {
      const int o = (i >> 2) | ((j >> 2) << 1) | ((k >> 2) << 2);
      // the ghost child of octant o, by register selects (no indexed private array)
      const int g = (o & 4) ? ((o & 2) ? ((o & 1) ? T.gch[7] : T.gch[6]) : ((o & 1) ? T.gch[5] : T.gch[4]))
                            : ((o & 2) ? ((o & 1) ? T.gch[3] : T.gch[2]) : ((o & 1) ? T.gch[1] : T.gch[0]));
      float sum = 0.f;
      if (g >= 0)
        for (int cid = 0; cid < 8; cid++)
          sum = sum + TMP[size_t(g) * 512 + ((2 * (i & 3) + (cid & 1)) * 8 + (2 * (j & 3) + ((cid >> 1) & 1))) * 8 +
                          (2 * (k & 3) + (cid >> 2))];
      out[k] = __builtin_fmaf(sum, 1.f, f[k]);
    }
  }
  store8(RHS + b, out);
}

// ProlongateAndUpdate128 (GMGSolver.cu:345-380) on the AMG hierarchy, natural
// order: GHOST x = parent x on INTERIOR cells, others x += pcoef parent x on
// INTERIOR cells, 0 elsewhere.
__global__ __launch_bounds__(64) void k_amg_prolong(Ctx c, const TileRec* __restrict__ list, int n, float* __restrict__ X,
                                                    AmgC A, float pcoef, int flags) {
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t b = size_t(T.self) * 512 + lane * 8;
  float x[8];
  load8(X + b, x);
  const uint64_t ct = load_ct8(A.ct + b);
  const int o = tr_octant(T);
  const size_t pb = size_t(T.parent) * 512 + (((o & 1) * 4 + (i >> 1)) * 8 + (((o >> 1) & 1) * 4 + (j >> 1))) * 8 +
                    ((o >> 2) & 1) * 4;
  float px[4];
  for (int q = 0; q < 4; q++) px[q] = X[pb + q];
  if (prob_done(c, T, flags)) return;
  const bool ghost = tr_type(T) == T_GHOST;
  for (int k = 0; k < 8; k++) {
    const bool in = ct_at(ct, k) & C_INT;
    const float cx = px[k >> 1];
    x[k] = in ? (ghost ? cx : __builtin_fmaf(pcoef, cx, x[k])) : 0.f;
  }
  store8(X + b, x);
}

// Fill whole tiles of a list with a value.
template <class V = float>
__global__ __launch_bounds__(64) void k_fill_tiles(const TileRec* __restrict__ list, int n, V* __restrict__ dst, float val) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  const size_t base = size_t(T.self) * 512 + lane_id() * 8;
  float v[8] = {val, val, val, val, val, val, val, val};
  store8(dst + base, v);
}
// dampedJacobiSolve pieces (GMGSolver.cu:479-525) on LEAF tiles:
// R = B - R on INTERIOR cells; then X += omega R / D (X = 0 where D == 0).
__global__ __launch_bounds__(64) void k_leaf_residual(Ctx c, const TileRec* __restrict__ list, int n,
                                                      const float* __restrict__ B, float* __restrict__ R) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  const size_t base = size_t(T.self) * 512 + lane_id() * 8;
  float b[8], r[8];
  load8(B + base, b);
  load8(R + base, r);
  const uint64_t ct = load_ct8(c.ctype + base);
  for (int k = 0; k < 8; k++) if (ct_at(ct, k) & C_INT) r[k] = b[k] - r[k];
  store8(R + base, r);
}
__global__ __launch_bounds__(64) void k_leaf_jacobi_update(Ctx c, const TileRec* __restrict__ list, int n, float* __restrict__ X,
                                                           const float* __restrict__ R, const float* __restrict__ D, float omega) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  const size_t base = size_t(T.self) * 512 + lane_id() * 8;
  float x[8], r[8], d[8];
  load8(X + base, x);
  load8(R + base, r);
  load8(D + base, d);
  const uint64_t ct = load_ct8(c.ctype + base);
  for (int k = 0; k < 8; k++)
    if (ct_at(ct, k) & C_INT) x[k] = (d[k] == 0.f) ? 0.f : x[k] + (omega * r[k]) / d[k];
  store8(X + base, x);
}

// ---------------------------------------------------------------------------
// Mixed-precision form: the FP64 outer loop of the iterative refinement (no
// reference counterpart).  x64 and the residual r64 are FP64 arrays beside the
// FP32 channels; the operator is FullNegativeLaplacian (GMGSolver.cu:203-213)
// evaluated in FP64 with the reference's per-face order, so the CPU
// transcription in FP64 gives the same bits.  Each inner solve is the FP32
// MGPCG on the FP32-rounded residual; its x is added to x64.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load8d(const double* __restrict__ p, double v[8]) {
  const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
  for (int h = 0; h < 4; h++) {
    const double2 t = q[h];
    v[2 * h] = t.x;
    v[2 * h + 1] = t.y;
  }
}
__device__ __forceinline__ void store8d(double* __restrict__ p, const double v[8]) {
  double2* q = reinterpret_cast<double2*>(p);
#pragma unroll
  for (int h = 0; h < 4; h++) q[h] = make_double2(v[2 * h], v[2 * h + 1]);
}
// Face coefficient before the NEUMANN cut in FP64 (GMGSolver.cu:26-39).
__device__ __forceinline__ double base_coeff64(uint8_t t0, uint8_t t1, double ooh) {
  const bool both = (t0 & T_LEAF) && (t1 & T_LEAF);
  const bool lg = ((t0 & T_LEAF) && (t1 & T_GHOST)) || ((t0 & T_GHOST) && (t1 & T_LEAF));
  return both ? ooh * ooh : (lg ? ((2.0 * ooh) / 3.0) * ooh : 0.0);
}

// Ghost x64 from the parent's INTERIOR value (PoissonGrid.cu:585-605).
__global__ __launch_bounds__(64) void k_propagate_ghosts64(Ctx c, const TileRec* __restrict__ list, int n, double* __restrict__ X) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t base = size_t(T.self) * 512 + lane * 8;
  const int oct = tr_octant(T), ox = oct & 1, oy = (oct >> 1) & 1, oz = (oct >> 2) & 1;
  const size_t pb = size_t(T.parent) * 512 + ((ox * 4 + (i >> 1)) << 6) + ((oy * 4 + (j >> 1)) << 3) + oz * 4;
  const double2 p0 = reinterpret_cast<const double2*>(X + pb)[0], p1 = reinterpret_cast<const double2*>(X + pb)[1];
  const double cv[4] = {p0.x, p0.y, p1.x, p1.y};
  const uint32_t pc = *reinterpret_cast<const uint32_t*>(c.ctype + pb);
  const uint64_t ct = load_ct8(c.ctype + base);
  double out[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const bool pin = uint8_t(pc >> (8 * (k >> 1))) & C_INT;
    out[k] = ((ct_at(ct, k) & C_INT) && pin) ? cv[k >> 1] : 0.0;
  }
  store8d(X + base, out);
}

// A x64 on every cell of the listed LEAF and GHOST tiles with their own types
// (NegativeLaplacianSameLevel128 with no subtree view, GMGSolver.cu:145-200),
// in FP64.  Neighbour pencils are read straight from memory (this runs once per
// refinement step, not per CG iteration).
__global__ __launch_bounds__(64) void k_apply64(Ctx c, const TileRec* __restrict__ list, int n, const double* __restrict__ X,
                                                double* __restrict__ AX) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  // the tile's pencils and its halo ring, index (i+1)*10 + (j+1), cells 0..7
  __shared__ double xs[100][8];
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t self = size_t(T.self) * 512;
  const size_t base = self + lane * 8;
  const double h = double(T.h);
  const double ooh = 1.0 / h;
  const uint8_t to = tr_type(T);
  const double* zp = zero_page<double>();
  // ---- loads, none in a branch: own pencil, one ring pencil (lanes 0..31;
  // x = 0 outside the grid), the two k halos, the face masks
  double own[8], ring[8];
  load8d(X + base, own);
  const int g = (lane >> 3) & 3, t = lane & 7;
  const int rn = pick4(g, uni(T.nbr[0]), uni(T.nbr[3]), uni(T.nbr[1]), uni(T.nbr[4]));
  const bool rok = lane < 32 && rn >= 0;
  load8d(rok ? X + size_t(rn) * 512 + ring_src(g, t) * 8 : zp, ring);
  const double xkm = *(T.nbr[2] >= 0 ? X + size_t(T.nbr[2]) * 512 + lane * 8 + 7 : zp);
  const double xkp = *(T.nbr[5] >= 0 ? X + size_t(T.nbr[5]) * 512 + lane * 8 : zp);
  const uint64_t m = load_ct8(own_masks(c, T) + lane * 8);
  double bf[6];
#pragma unroll
  for (int f = 0; f < 6; f++) bf[f] = base_coeff64(to, T.nbr[f] < 0 ? to : tr_ntype(T, f), ooh);
  const double bint = base_coeff64(to, to, ooh);
  const int pc = (i + 1) * 10 + (j + 1);
#pragma unroll
  for (int k = 0; k < 8; k++) xs[pc][k] = own[k];
  if (lane < 32) {
#pragma unroll
    for (int k = 0; k < 8; k++) xs[ring_dst(g, t)][k] = ring[k];
  }
  wave_sync();
  const double bim = (i == 0) ? bf[0] : bint, bip = (i == 7) ? bf[3] : bint;
  const double bjm = (j == 0) ? bf[1] : bint, bjp = (j == 7) ? bf[4] : bint;
  double out[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const double xn[6] = {xs[pc - 10][k], xs[pc + 10][k], xs[pc - 1][k], xs[pc + 1][k], k == 0 ? xkm : own[k - 1],
                          k == 7 ? xkp : own[k + 1]};
    const double bc[6] = {bim, bip, bjm, bjp, k == 0 ? bf[2] : bint, k == 7 ? bf[5] : bint};
    const uint32_t mk = ct_at(m, k);
    double sum = 0.0;
#pragma unroll
    for (int e = 0; e < 6; e++) sum = __builtin_fma((mk & (1u << e)) ? bc[e] : 0.0, own[k] - xn[e], sum);
    out[k] = (mk & M_INT) ? ((sum * h) * h) * h : 0.0;
  }
  store8d(AX + base, out);
}

// Ghost rows of A x64 summed into the parent (AccumulateToParents128 with
// coefficient 1, additive, all cell types; PoissonGrid.cu:664-719).
__global__ __launch_bounds__(64) void k_accumulate_ghosts64(Ctx c, const TileRec* __restrict__ list, int n, double* __restrict__ AX) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id();
  const size_t base = size_t(T.self) * 512 + lane * 8;
  __shared__ double gl[512];
  double v[8];
  load8d(AX + base, v);
  const uint64_t ct = load_ct8(c.ctype + base);
  for (int k = 0; k < 8; k++) gl[lane * 8 + k] = (ct_at(ct, k) & (C_INT | C_DIR | C_NEU)) ? v[k] : 0.0;
  wave_sync();
  const int a = lane >> 4, bq = (lane >> 2) & 3, cq = lane & 3;
  double sum = 0.0;
  for (int ii = 0; ii < 2; ii++)
    for (int jj = 0; jj < 2; jj++)
      for (int kk = 0; kk < 2; kk++) sum += gl[((2 * a + ii) * 8 + (2 * bq + jj)) * 8 + 2 * cq + kk];
  const size_t pcell = parent_cell(T, lane);
  AX[pcell] = AX[pcell] + sum * 1.0;
}

// |b|^2 partials of the LEAF tiles (per leaf slot).
__global__ __launch_bounds__(64) void k_ir_norm(Ctx c, const TileRec* __restrict__ list, int n, const float* __restrict__ B) {
  const int flags = F_SWIZZLE;
  TILE_KERNEL_PROLOGUE
  const size_t base = size_t(T.self) * 512 + lane_id() * 8;
  float b[8];
  load8(B + base, b);
  double a = 0.0;
  for (int k = 0; k < 8; k++) a += double(b[k]) * double(b[k]);
  a = wave_sum(a);
  if (lane_id() == 0) c.part[T.leaf_slot] = a;
}
// X *= 2^(sign * scale_exp) on every cell of the LEAF tiles (exact: a power of two).
__global__ __launch_bounds__(64) void k_scale_leaf(Ctx c, const TileRec* __restrict__ list, int n, float* __restrict__ X,
                                                   int sign) {
  const int flags = F_SWIZZLE;
  TILE_KERNEL_PROLOGUE
  const float f = ldexpf(1.f, sign * c.ps[T.prob].scale_exp);
  const size_t base = size_t(T.self) * 512 + lane_id() * 8;
  float v[8];
  load8(X + base, v);
  for (int k = 0; k < 8; k++) v[k] = v[k] * f;
  store8(X + base, v);
}
// r64 = b - A x64 on every cell of the LEAF tiles (R holds A x64 on entry).
//   mode 0: the next inner solve's rhs r = float(r64), partials of |r64|^2;
//   mode 1 (pure Neumann, before re-centring): partials of sum r64 and of the
//           INTERIOR count (MeanAsync, PoissonGrid.cu:231-325);
//   mode 2: partials of |r64|^2 only (a check of a finished solve).
__global__ __launch_bounds__(64) void k_ir_residual(Ctx c, const TileRec* __restrict__ list, int n, const float* __restrict__ B,
                                                    double* __restrict__ R, int mode) {
  const int flags = F_SWIZZLE;
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id();
  const size_t base = size_t(T.self) * 512 + lane * 8;
  float b[8];
  double r[8];
  load8(B + base, b);
  load8d(R + base, r);
  for (int k = 0; k < 8; k++) r[k] = double(b[k]) - r[k];
  store8d(R + base, r);
  double a = 0.0, q = 0.0;
  if (mode == 0) {
    float rf[8];
    for (int k = 0; k < 8; k++) { rf[k] = float(r[k]); a += r[k] * r[k]; }
    store8(c.r + base, rf);
  } else if (mode == 2) {
    for (int k = 0; k < 8; k++) a += r[k] * r[k];
  } else {
    const uint64_t ct = load_ct8(c.ctype + base);
    for (int k = 0; k < 8; k++) { a += r[k]; q += (ct_at(ct, k) & C_INT) ? 1.0 : 0.0; }
    q = wave_sum(q);
  }
  a = wave_sum(a);
  if (lane == 0) {
    c.part[T.leaf_slot] = a;
    if (mode == 1) c.part2[T.leaf_slot] = q;
  }
}
// Pure Neumann: r64 -= mean on INTERIOR cells (ReCenterLeafVoxels,
// PoissonSolver.cu:356-368, in FP64), then as k_ir_residual mode 0.
__global__ __launch_bounds__(64) void k_ir_recenter(Ctx c, const TileRec* __restrict__ list, int n, double* __restrict__ R,
                                                    int write_r) {
  const int flags = F_SWIZZLE;
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id();
  const size_t base = size_t(T.self) * 512 + lane * 8;
  const double mean = c.ps[T.prob].mean;
  double r[8];
  load8d(R + base, r);
  const uint64_t ct = load_ct8(c.ctype + base);
  for (int k = 0; k < 8; k++) if (ct_at(ct, k) & C_INT) r[k] = r[k] - mean;
  store8d(R + base, r);
  float rf[8];
  double a = 0.0;
  for (int k = 0; k < 8; k++) { rf[k] = float(r[k]); a += r[k] * r[k]; }
  if (write_r) store8(c.r + base, rf);
  a = wave_sum(a);
  if (lane == 0) c.part[T.leaf_slot] = a;
}
// x64 = x on LEAF tiles (the FP32 pressure of a finished solve, to be checked).
__global__ __launch_bounds__(64) void k_ir_load_x(Ctx c, const TileRec* __restrict__ list, int n, const float* __restrict__ X,
                                                  double* __restrict__ X64) {
  const int flags = F_SWIZZLE;
  TILE_KERNEL_PROLOGUE
  const size_t base = size_t(T.self) * 512 + lane_id() * 8;
  float xf[8];
  double xd[8];
  load8(X + base, xf);
  for (int k = 0; k < 8; k++) xd[k] = double(xf[k]);
  store8d(X64 + base, xd);
}
// x64 += x (the inner solve's correction) on LEAF tiles; finish: x = float(x64).
__global__ __launch_bounds__(64) void k_ir_accumulate(Ctx c, const TileRec* __restrict__ list, int n, double* __restrict__ X64,
                                                      int finish) {
  const int flags = F_SWIZZLE;
  TILE_KERNEL_PROLOGUE
  const size_t base = size_t(T.self) * 512 + lane_id() * 8;
  double xd[8];
  float xf[8];
  load8d(X64 + base, xd);
  if (finish) {
    for (int k = 0; k < 8; k++) xf[k] = float(xd[k]);
    store8(c.x + base, xf);
    return;
  }
  load8(c.x + base, xf);
  for (int k = 0; k < 8; k++) xd[k] = xd[k] + double(xf[k]);
  store8d(X64 + base, xd);
}
// Per problem q (workgroup -> problem through pmap, else the identity): its
// |r64|^2 (the leaf-slot partials) and |b|^2 (ProbState::bb_b) into
// out[2q], out[2q + 1].
__global__ __launch_bounds__(FIN_THREADS) void k_ir_verdict(Ctx c, double* __restrict__ out) {
  __shared__ double lds[FIN_CHUNK];
  const int q = c.pmap ? c.pmap[blockIdx.x] : int(blockIdx.x);
  const ProbState* ps = c.ps + q;
  const double s = block_sum_partials(c.part, ps->leaf_begin, ps->leaf_count, lds);
  if (threadIdx.x == 0) {
    out[2 * q] = s;
    out[2 * q + 1] = ps->bb_b;
  }
}
// Per-problem totals of the leaf-slot partials, in the fixed order.
__global__ __launch_bounds__(FIN_THREADS) void k_prob_sums(Ctx c, double* __restrict__ out) {
  __shared__ double lds[FIN_CHUNK];
  const ProbState* ps = c.ps + blockIdx.x;
  const double s = block_sum_partials(c.part, ps->leaf_begin, ps->leaf_count, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// ---------------------------------------------------------------------------
// The projection around the solve (FluidEuler::project, FluidEuler.h:581-612):
// face velocities u, v, w (one value per cell: its minus face, natural
// layout), their divergence as the right-hand side, and the pressure gradient
// added back.  One wavefront per tile, lane = pencil; neighbour pencils are
// read straight from memory (these run once per projection).
// ---------------------------------------------------------------------------
struct Vel {
  float* u[3];
};
__device__ __forceinline__ void load8c(const uint8_t* __restrict__ p, uint8_t v[8]) {
  const uint64_t w = load_ct8(p);
#pragma unroll
  for (int k = 0; k < 8; k++) v[k] = ct_at(w, k);
}

// Ghost velocities from the parents (PropagateToChildren with cell types
// INTERIOR|DIRICHLET|NEUMANN, PoissonGrid.cu:635-661): a typed ghost cell takes
// its parent cell's value (0 if the parent cell is untyped).
__global__ __launch_bounds__(64) void k_propagate_vel(Ctx c, const TileRec* __restrict__ list, int n, Vel U) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t base = size_t(T.self) * 512 + lane * 8;
  const int oct = tr_octant(T), ox = oct & 1, oy = (oct >> 1) & 1, oz = (oct >> 2) & 1;
  const size_t pb = size_t(T.parent) * 512 + ((ox * 4 + (i >> 1)) << 6) + ((oy * 4 + (j >> 1)) << 3) + oz * 4;
  const uint32_t pc = *reinterpret_cast<const uint32_t*>(c.ctype + pb);
  const uint64_t ct = load_ct8(c.ctype + base);
  const uint8_t mask = C_INT | C_DIR | C_NEU;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float4 pv = load4(U.u[a] + pb);
    const float cv[4] = {pv.x, pv.y, pv.z, pv.w};
    float v[8];
    load8(U.u[a] + base, v);
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (!(ct_at(ct, k) & mask)) continue;
      const bool pin = uint8_t(pc >> (8 * (k >> 1))) & mask;
      v[k] = pin ? cv[k >> 1] : 0.f;
    }
    store8(U.u[a] + base, v);
  }
}

// Does a face between tiles of types t0 (the cell's) and t1 count?  LEAF-LEAF,
// LEAF-GHOST, GHOST-LEAF (PoissonSolver.cu:113-119).
__device__ __forceinline__ bool flux_face(uint8_t t0, uint8_t t1) {
  return (t0 == T_LEAF && (t1 == T_LEAF || t1 == T_GHOST)) || (t0 == T_GHOST && t1 == T_LEAF);
}

// VelocityVolumeDivergenceOnLeafs (PoissonSolver.cu:66-137) on LEAF|GHOST
// tiles: INTERIOR cells sum u h^2 over their counted faces in the order x-,
// x+, y-, y+, z-, z+ (-u of the own minus face, +u of the plus neighbour's,
// 0 outside the grid, where the neighbour counts as the tile's own type); the
// other cells get 0.
__global__ __launch_bounds__(64) void k_divergence(Ctx c, const TileRec* __restrict__ list, int n, Vel U,
                                                   float* __restrict__ DIV) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t self = size_t(T.self) * 512;
  const size_t base = self + lane * 8;
  const uint8_t t0 = tr_type(T);
  const float h = T.h;
  float own[3][8];
#pragma unroll
  for (int a = 0; a < 3; a++) load8(U.u[a] + base, own[a]);
  // plus neighbours: u of pencil (i+1,j), v of pencil (i,j+1), w of cell (i,j,k+1)
  float up[8], vp[8], wkp = 0.f;
  uint8_t tx = t0, ty = t0, tz = t0;  // tile types across the plus faces on the tile boundary
  {
    const int nx = T.nbr[3], ny = T.nbr[4], nz = T.nbr[5];
    if (i < 7) load8(U.u[0] + base + 64, up);
    else if (nx >= 0) load8(U.u[0] + size_t(nx) * 512 + j * 8, up);
    else for (int k = 0; k < 8; k++) up[k] = 0.f;
    if (j < 7) load8(U.u[1] + base + 8, vp);
    else if (ny >= 0) load8(U.u[1] + size_t(ny) * 512 + i * 64, vp);
    else for (int k = 0; k < 8; k++) vp[k] = 0.f;
    if (nz >= 0) wkp = U.u[2][size_t(nz) * 512 + lane * 8];
    if (i == 7 && nx >= 0) tx = tr_ntype(T, 3);
    if (j == 7 && ny >= 0) ty = tr_ntype(T, 4);
    if (nz >= 0) tz = tr_ntype(T, 5);
  }
  // tile types across the minus faces on the tile boundary
  const uint8_t sx = (i == 0 && T.nbr[0] >= 0) ? tr_ntype(T, 0) : t0;
  const uint8_t sy = (j == 0 && T.nbr[1] >= 0) ? tr_ntype(T, 1) : t0;
  const uint8_t sz = (T.nbr[2] >= 0) ? tr_ntype(T, 2) : t0;
  const uint64_t ct = load_ct8(c.ctype + base);
  float out[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    float sum = 0.f;
    if (ct_at(ct, k) & C_INT) {
      const float fu[6] = {-own[0][k], up[k], -own[1][k], vp[k], -own[2][k], k < 7 ? own[2][k + 1] : wkp};
      const uint8_t ft[6] = {sx, tx, sy, ty, k == 0 ? sz : t0, k == 7 ? tz : t0};
#pragma unroll
      for (int e = 0; e < 6; e++)
        if (flux_face(t0, ft[e])) sum = __builtin_fmaf(fu[e] * h, h, sum);
    }
    out[k] = sum;
  }
  store8(DIV + base, out);
}

// AccumulateToParents(ch, ch, -1, LEAF, SUBTREE, mask, coeff, additive)
// (PoissonGrid.cu:721-755) on LEAF tiles: each cell whose type is in mask adds
// coeff times the sum, in child order (x fastest), of its 8 children (cells of
// its ghost child tile) whose type is in mask.
__device__ __forceinline__ void tile_accumulate_children(const Ctx& c, const TileRec& T, float* __restrict__ X,
                                                         uint8_t mask, float coeff) {
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t base = size_t(T.self) * 512 + lane * 8;
  const uint64_t ct = load_ct8(c.ctype + base);
  float x[8];
  load8(X + base, x);
  // ghost children of the pencil's two octants (k 0..3, k 4..7)
  const int o0 = (i >> 2) | ((j >> 2) << 1);
  const int g[2] = {pick4(o0, uni(T.gch[0]), uni(T.gch[1]), uni(T.gch[2]), uni(T.gch[3])),
                    pick4(o0, uni(T.gch[4]), uni(T.gch[5]), uni(T.gch[6]), uni(T.gch[7]))};
  float fv[2][2][2][8];  // [oz][ii][jj][fine k]
  uint8_t fc[2][2][2][8];
#pragma unroll
  for (int oz = 0; oz < 2; oz++)
#pragma unroll
    for (int ii = 0; ii < 2; ii++)
#pragma unroll
      for (int jj = 0; jj < 2; jj++) {
        const int fi = 2 * (i & 3) + ii, fj = 2 * (j & 3) + jj;
        if (g[oz] >= 0) {
          const size_t fb = size_t(g[oz]) * 512 + (fi * 8 + fj) * 8;
          load8(X + fb, fv[oz][ii][jj]);
          load8c(c.ctype + fb, fc[oz][ii][jj]);
        } else {
#pragma unroll
          for (int k = 0; k < 8; k++) { fv[oz][ii][jj][k] = 0.f; fc[oz][ii][jj][k] = 0; }
        }
      }
#pragma unroll
  for (int k = 0; k < 8; k++) {
    if (!(ct_at(ct, k) & mask)) continue;
    const int oz = k >> 2, kl = 2 * (k & 3);
    float sum = 0.f;
#pragma unroll
    for (int cid = 0; cid < 8; cid++) {
      const int ii = cid & 1, jj = (cid >> 1) & 1, kk = (cid >> 2) & 1;
      if (fc[oz][ii][jj][kl + kk] & mask) sum += fv[oz][ii][jj][kl + kk];
    }
    x[k] = x[k] + sum * coeff;
  }
  store8(X + base, x);
}
__global__ __launch_bounds__(64) void k_accumulate_children(Ctx c, const TileRec* __restrict__ list, int n,
                                                            float* __restrict__ X, uint8_t mask, float coeff) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  tile_accumulate_children(c, T, X, mask, coeff);
}
// The same for the three face-velocity channels in one launch (one after the
// other inside each wavefront).
__global__ __launch_bounds__(64) void k_accumulate_children_vel(Ctx c, const TileRec* __restrict__ list, int n, Vel U,
                                                                uint8_t mask, float coeff) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  for (int a = 0; a < 3; a++) tile_accumulate_children(c, T, U.u[a], mask, coeff);
}

// AddGradientToFaceCenters (PoissonSolver.cu:144-204) on LEAF|GHOST tiles
// (ghost p already taken from the parents): ghost velocities start at 0; each
// cell's minus face of each axis gets u += (p1 - p0) / dh with p1, p0 the
// INTERIOR values (0 elsewhere) of the cell and its minus neighbour (0 outside
// the grid, where the neighbour counts as the tile's own type and the cell's
// own cell type), dh = h across LEAF-LEAF faces and 1.5 h across LEAF-GHOST /
// GHOST-LEAF faces, no update across other faces or where either cell is NEUMANN.
__global__ __launch_bounds__(64) void k_gradient(Ctx c, const TileRec* __restrict__ list, int n, const float* __restrict__ P,
                                                 Vel U) {
  const int flags = 0;
  TILE_KERNEL_PROLOGUE
  const int lane = lane_id(), i = lane >> 3, j = lane & 7;
  const size_t self = size_t(T.self) * 512;
  const size_t base = self + lane * 8;
  const uint8_t t0 = tr_type(T);
  const float h = T.h;
  const float dh_lg = float(1.5 * double(h));
  uint8_t c0[8];
  float p1[8];
  load8c(c.ctype + base, c0);
  load8(P + base, p1);
  // minus neighbours: pencil (i-1,j), pencil (i,j-1), cell (i,j,-1)
  float px[8], py[8], pz = 0.f;
  uint8_t cx[8], cy[8], cz = 0;
  bool ex = false, ey = false, ez = false;  // outside the grid
  {
    const int nx = T.nbr[0], ny = T.nbr[1], nz = T.nbr[2];
    if (i > 0) { load8(P + base - 64, px); load8c(c.ctype + base - 64, cx); }
    else if (nx >= 0) { load8(P + size_t(nx) * 512 + (56 + j) * 8, px); load8c(c.ctype + size_t(nx) * 512 + (56 + j) * 8, cx); }
    else ex = true;
    if (j > 0) { load8(P + base - 8, py); load8c(c.ctype + base - 8, cy); }
    else if (ny >= 0) { load8(P + size_t(ny) * 512 + (i * 8 + 7) * 8, py); load8c(c.ctype + size_t(ny) * 512 + (i * 8 + 7) * 8, cy); }
    else ey = true;
    if (nz >= 0) { pz = P[size_t(nz) * 512 + lane * 8 + 7]; cz = c.ctype[size_t(nz) * 512 + lane * 8 + 7]; }
    else ez = true;
  }
  const uint8_t sx = (i == 0 && T.nbr[0] >= 0) ? tr_ntype(T, 0) : t0;
  const uint8_t sy = (j == 0 && T.nbr[1] >= 0) ? tr_ntype(T, 1) : t0;
  const uint8_t sz = (T.nbr[2] >= 0) ? tr_ntype(T, 2) : t0;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float u[8];
    if (t0 == T_GHOST) { for (int k = 0; k < 8; k++) u[k] = 0.f; }
    else load8(U.u[a] + base, u);
#pragma unroll
    for (int k = 0; k < 8; k++) {
      bool outside;
      uint8_t nt, nc;
      float pn;
      if (a == 0) { outside = (i == 0) && ex; nt = sx; nc = outside ? c0[k] : cx[k]; pn = outside ? 0.f : px[k]; }
      else if (a == 1) { outside = (j == 0) && ey; nt = sy; nc = outside ? c0[k] : cy[k]; pn = outside ? 0.f : py[k]; }
      else if (k > 0) { outside = false; nt = t0; nc = c0[k - 1]; pn = p1[k - 1]; }
      else { outside = ez; nt = sz; nc = outside ? c0[0] : cz; pn = outside ? 0.f : pz; }
      const float pa = (c0[k] & C_INT) ? p1[k] : 0.f;
      const float pb = (nc & C_INT) ? pn : 0.f;
      const bool ll = (t0 & T_LEAF) && (nt & T_LEAF);
      const bool lg = ((t0 & T_LEAF) && (nt & T_GHOST)) || ((t0 & T_GHOST) && (nt & T_LEAF));
      if (!(ll || lg) || (c0[k] & C_NEU) || (nc & C_NEU)) continue;
      u[k] = u[k] + (pa - pb) / (ll ? h : dh_lg);
    }
    store8(U.u[a] + base, u);
  }
}

}  // namespace cg
