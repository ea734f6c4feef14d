// librqx: gfx950 kernels for the SQD subspace Pauli-sum action.
//
// Index layout (built once per plan, resident in HBM):
//   keys[N]      sorted unique uint64 states (bit q = qubit q)
//   segments     the states split by their high part s >> L: seg_hi[M] (sorted),
//                seg_off[M+1] row ranges; lo[N] = s & (2^L - 1) as uint32
//   classes      the X groups split by xmask >> L (groups are sorted by xmask,
//                so each class is a contiguous group range cls_goff[C+1])
//   tiles        <= 256 consecutive rows of one segment (one workgroup each)
//
// For a tile of segment m and a class c, every row's flipped state s ^ x has
// the same high part seg_hi[m] ^ cls_hi[c]: the class's target segment is
// looked up once per (tile, class), and when it exists its low keys are staged
// in LDS and each row runs a branchless binary search over them for each group
// of the class.  Misses at the segment level resolve whole (tile, class)
// blocks of (row, group) pairs with one lookup.
//
// Compile: hipcc --offload-arch=gfx950 -O3 -fPIC -shared (see rqutils_amd/_build.py)

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <initializer_list>
#include <type_traits>
#include <utility>
#include <vector>

#include "rqx.h"

namespace {

constexpr int TILE = 256;          // rows per workgroup (4 waves of 64)
constexpr int SEG_CAP = 8192;      // max rows of a segment (LDS: 32 KiB of uint32 keys)
constexpr int NUM_XCD = 8;         // MI355X accelerator complex dies

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(RQX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

enum { MODE_MATVEC = 0, MODE_LOOKUP = 1, MODE_COUNT = 2, MODE_GRAD = 3, MODE_ROWCOUNT = 4, MODE_FILL = 5, MODE_COO = 6,
       MODE_FILLCOL = 7 };

// hit cache: 64 rows per slice (one wave), two sections per slice.  Section A
// holds the hits of groups that flip no bit below the problem's section split
// (in a product subspace: the source is the row's own low string in another
// high block, so every row of a slice hits the same such groups and the level
// is a row offset with, mostly, one value); section B the other hits, stored as
// SELL-8-64 (below).  The column pass first writes every hit into a full layout
// of packets of PACK consecutive entries of one row: entry k of row r (k
// counted over section A, then B; both widths multiples of PACK) at
//   off[s] + (k / PACK) 64 PACK + (q / e) 64 e + (r % 64) e + q % e,
// q = k % PACK, s = r / 64, e = the entries of the array's type in 16 bytes (4
// for columns and FP32 values, 2 for FP64 or complex FP32, 1 for complex FP64),
// so each 16-byte access of a wave covers 1 KB contiguously; the stored layout
// keeps from it the columns that are not row offsets.  A row's y is (diagonal
// + B hits) + A hits, each part accumulated in group order, in the matrix-free
// kernel too.
constexpr int SLICE = 64;
constexpr int PACK = 4;

template <typename T> constexpr int pack_sub() { return (16 / sizeof(T)) < PACK ? static_cast<int>(16 / sizeof(T)) : PACK; }

template <typename T> __host__ __device__ inline int64_t packed_slot(int64_t off, uint32_t k, int lane) {
    constexpr int E = pack_sub<T>();
    const uint32_t q = k % PACK;
    return off + static_cast<int64_t>(k / PACK) * (SLICE * PACK) + static_cast<int64_t>(q / E) * (SLICE * E) +
           lane * E + (q % E);
}

typedef int32_t i32x4 __attribute__((ext_vector_type(4)));

// the packet of entries kb .. kb + PACK - 1 (kb a multiple of PACK) of a row, as
// 16-byte stores
template <typename T> __device__ inline void store_packet(void *a, int64_t off, uint32_t kb, int lane, const T *v) {
    constexpr int E = pack_sub<T>();
#pragma unroll
    for (int i = 0; i < PACK / E; ++i) {
        i32x4 t;
        __builtin_memcpy(&t, v + i * E, 16);
        *reinterpret_cast<i32x4 *>(static_cast<T *>(a) + packed_slot<T>(off, kb + i * E, lane)) = t;
    }
}

// Section B is stored as SELL-8-64: the rows of a slice in descending order of
// their section-B entry count (position l = 0..63, ties in row order), chunk j =
// positions [8 j, 8 j + 8) padded to its widest row, W[j] packets.  Packet level q
// holds the packets of the chunks with W[j] > q -- a prefix of the positions,
// since the widths descend -- lane-contiguous: entry k (q = k / PACK, u = k % PACK)
// of position l sits at
//   off[s] + 32 sum_j min(q, W[j]) + (u / e) n_q e + l e + u % e,   n_q = 8 #{j : W[j] > q}.
constexpr int CH = 8;                  // rows per chunk
constexpr int NCH = SLICE / CH;        // chunks per slice

struct SellLevel {
    int64_t base;   // first entry of the level, from the slice's first entry
    int32_t n;      // positions holding a packet at this level
};

__host__ __device__ inline SellLevel sell_level(const uint32_t *W, uint32_t q) {
    int64_t b = 0;
    int32_t m = 0;
#pragma unroll
    for (int j = 0; j < NCH; ++j) {
        b += W[j] < q ? W[j] : q;
        m += W[j] > q ? 1 : 0;
    }
    return SellLevel{b * (CH * PACK), m * CH};
}

template <typename T> __host__ __device__ inline int64_t sell_slot(int64_t off, const SellLevel &lv, uint32_t u, int pos) {
    constexpr int E = pack_sub<T>();
    return off + lv.base + static_cast<int64_t>(u / E) * (lv.n * E) + pos * E + (u % E);
}

// the packet of level q of position pos, as 16-byte stores
template <typename T>
__device__ inline void store_packet_sell(void *a, int64_t off, const uint32_t *W, uint32_t q, int pos, const T *v) {
    constexpr int E = pack_sub<T>();
    const SellLevel lv = sell_level(W, q);
#pragma unroll
    for (int i = 0; i < PACK / E; ++i) {
        i32x4 t;
        __builtin_memcpy(&t, v + i * E, 16);
        *reinterpret_cast<i32x4 *>(static_cast<T *>(a) + sell_slot<T>(off, lv, i * E, pos)) = t;
    }
}

// the chunk widths of a slice (8 x uint16 in 16 bytes), widened
__device__ inline void load_chunk_widths(const uint16_t *bw, int64_t slice, uint32_t *W) {
    const i32x4 t = *reinterpret_cast<const i32x4 *>(bw + slice * NCH);
#pragma unroll
    for (int j = 0; j < NCH / 2; ++j) {
        W[2 * j] = static_cast<uint32_t>(t[j]) & 0xffffu;
        W[2 * j + 1] = static_cast<uint32_t>(t[j]) >> 16;
    }
}

// column / value pass: the last PACK section-B entries of a row (column, or
// value and, when both precisions are filled, FP32 value), stored once the
// packet is whole
template <typename FV, typename FL> struct FillBuf {
    int32_t c[PACK];
    FV v[PACK];
    FL l[PACK];
};

// terms fetched per scalar-memory round trip when a coefficient sum is evaluated;
// the z and coefficient arrays are padded by as many entries
constexpr int TERM_PAD = 4;

// A load of plan data (term and group tables, target lists, tile lists) through
// the constant address space: the kernels never write these arrays, and saying
// so lets the compiler serve wave-uniform loads from the scalar cache even in
// kernels that store inside their loops (the cache passes, the gradient), where
// it must otherwise assume a store may have changed them and load with vector
// instructions.
template <typename T> __device__ inline T ldc(const T *a) {
    if constexpr (std::is_arithmetic<T>::value) {
        return *(const __attribute__((address_space(4))) T *)(a);
    } else {                      // structs and vector types: word by word
        static_assert(sizeof(T) % 4 == 0, "whole words");
        const __attribute__((address_space(4))) uint32_t *w = (const __attribute__((address_space(4))) uint32_t *)(a);
        uint32_t t[sizeof(T) / 4];
#pragma unroll
        for (size_t k = 0; k < sizeof(T) / 4; ++k) t[k] = w[k];
        T v;
        __builtin_memcpy(&v, t, sizeof(T));
        return v;
    }
}

// One problem of a plan (a batched plan holds several).  Row, segment, class,
// group and term indices in the plan arrays are absolute; the descriptor gives
// the problem's own ranges.
struct ProbDesc {
    uint32_t seg_base, n_segs;      // segments [seg_base, seg_base + n_segs)
    uint32_t cls_base, n_classes;   // classes  [cls_base, cls_base + n_classes)
    uint32_t grp_base;              // first group (the x == 0 group when has_diag)
    uint32_t lomask;                // low-key mask 2^L - 1 of the problem's split
    int32_t has_diag;
    int32_t a_bits;                 // section split: groups with x & (2^a_bits - 1) == 0 are section A (a_bits >= L)
    uint32_t row0;                  // the problem's first row
    uint32_t al_nlo;                // alpha lists: low strings n_lo of the product (rows b * n_lo + a), or 0
    uint32_t al_a0;                 //   the problem's first alpha index in the list arrays (a multiple of 64)
    uint32_t prod;                  // 1: the states are a product at a_bits (a_bits > L)
    uint32_t al_nhi;                // alpha lists: high strings of the product (rows / n_lo)
    uint32_t pm_base;               // partner map: first element (UINT32_MAX: none), its row length
    uint32_t pm_na;                 //   (alpha patterns), the problem's first beta index of the
    uint32_t bp_b0;                 //   beta-partner offsets
    uint32_t bl_b0;                 // beta lists: the problem's first beta index of their offsets
    uint32_t full;                  // 1: lists cover every group (no class is searched)
    uint32_t al_g0, al_gnr;         // alpha lists' G tables: first element, values per beta block
    uint32_t al_pat0;               // partner map: the problem's first alpha pattern in the plan's list,
    uint32_t al_hsl;                //   log2 of the table values per pattern the product kernel stages
    uint32_t pad0, pad1;
};

__device__ inline uint64_t section_mask(int32_t a_bits) { return a_bits >= 64 ? ~0ull : ((1ull << a_bits) - 1ull); }

// What the coefficient evaluator needs of a group, fetched with one scalar load
// (separate arrays cost one dependent round trip each: scalar loads return out
// of order, so every use waits for all loads in flight)
struct __attribute__((aligned(32))) GroupDesc {
    uint64_t x;          // X mask
    uint64_t z0;         // Z mask of the group's first term
    uint32_t k0, k1;     // its terms [k0, k1)
    int32_t tab;         // its sign table, or -1
    uint32_t xpos;       // with a sign table: the positions of x's bits, 6 bits each from
                         // the lowest, and their count in bits 24..31
};

// Alpha lists, the matvec of an exact product subspace (rows b * n_lo + a with
// key(b * n_lo + a) = beta_b << a_bits | alpha_a): the groups that flip only
// alpha bits and whose terms' Z masks hold only alpha bits hit row i = b * n_lo + a
// at b * n_lo + a' with a value that depends on alpha_a alone.  Per alpha index a,
// its hits of those groups in group order -- a', the value in the matvec's
// precision and (a side array) the group -- replace the searches of all those
// groups for every row.  Entry k of alpha index a sits at boff[a >> 6] + 64 k +
// (a & 63): the lanes of a wave (consecutive alpha indices) read consecutive
// entries.  The beta lists use the same entries with the group in place of the
// split word.
template <typename R> struct AlEnt;
template <> struct __attribute__((aligned(16))) AlEnt<float> {
    union {
        uint32_t sw;     // alpha lists: the split word (0: a plain value)
        uint32_t g;      // beta lists: the group
    };
    uint32_t col;
    float2 v;
};
template <> struct __attribute__((aligned(16))) AlEnt<double> {
    union {
        uint32_t sw;
        uint32_t g;
    };
    uint32_t col, pad0, pad1;
    double2 v;
};
// An entry whose column carries AL_SPLIT is a group whose terms also hold Z bits
// of the other side, split as C(s) = F + sum_j (-1)^popc(s & zeta_j) G_j: zeta_j
// the distinct parts on the entry's side of its terms with the other side's Z
// bits, G_j (a function of the other side's string) the sum of those terms'
// coefficients times (-1)^popc(s & z_k) over the other side's bits, F the sum over
// its terms without the other side's Z bits (each sum in term order).  The
// entry's value is F.
//  - alpha lists: the problem's G tables hold, per beta block b, one row of
//    al_gnr values at al_g0 + b * al_gnr (a beta block's G values are contiguous,
//    and the product kernel stages them in LDS); the split word holds the
//    group's first row in bits 0..27, its pattern count in bits 28..29 and the
//    signs (-1)^popc(alpha_a & zeta_j) of the entry's alpha string in bit 30 + j
//  - beta lists: the group's SplitDesc gives the zeta_j and its rows of the G
//    tables (one row of n_lo values per j, over the alpha strings)
constexpr uint32_t AL_SPLIT = 0x80000000u;
constexpr int SPLIT_MAX = 2;
constexpr uint32_t SW_ROW_BITS = 28;
// The alpha-list entries once more for real coefficients (what the product kernel
// reads then): one word -- a' in bits 0..13, the split word's pattern count in
// bits 14..15, its signs in bits 16..17, its first G row in bits 18..31 -- and the
// real part of the value.  (The product kernel runs only where a beta block's x and
// G values fit in 64 KiB of LDS, so a' and the row stay below 2^14 there.)
constexpr uint32_t CE_BITS = 14;
template <typename R> struct AlEntC;
template <> struct __attribute__((aligned(8))) AlEntC<float> {
    uint32_t w;
    float v;
};
template <> struct __attribute__((aligned(16))) AlEntC<double> {
    uint32_t w, pad;
    double v;
};
struct __attribute__((aligned(16))) SplitDesc {
    uint64_t zeta[SPLIT_MAX];
    uint32_t grow;                  // first row of the G tables (alpha groups: of the problem's rows)
    uint32_t n_p;                   // patterns (<= SPLIT_MAX; 0: F is the whole sum)
    uint32_t split;                 // 1: the group's entries are split, 0: plain values
    uint32_t pad0;
};

// One (class, target segment) pair of a segment's target list, with what the
// tile kernel needs of both (one scalar load)
struct __attribute__((aligned(32))) TargetDesc {
    uint32_t ta, tn;     // the target segment's rows [ta, ta + tn)
    uint32_t ho, lg;     // its hash table (hashed plans): first entry, log2 buckets
    uint32_t g0, g1;     // the class's groups [g0, g1), without the x == 0 group
    uint32_t cls, ts;    // the class and the target segment
};

struct KParams {
    const uint64_t *keys;
    const uint32_t *lo;
    const uint64_t *seg_hi;
    const uint32_t *seg_off;
    const uint64_t *cls_hi;
    const uint32_t *cls_goff;
    const uint32_t *cls_lane;      // per class: bit 0 when every group lane_evaluable takes (the matvec's
                                   //   fast loop), bit 1 when its groups flip only alpha bits (alpha lists)
    const uint32_t *st_off;        // per segment: its first (class, target segment) pair
    const uint32_t *h_off;         // hashed segments (NULL: binary search): first table entry,
    const uint32_t *h_log;         //   log2 of the bucket count,
    const uint32_t *h_tab;         //   tables of (low key << 13 | rank), two entries per bucket
    const TargetDesc *st_list;     // pairs of the classes whose target segment exists, class order
    const uint32_t *g_xlo;
    const uint32_t *g_toff;
    const uint64_t *z;
    const void *coef;      // double2[T] or float2[T]
    const int32_t *g_tab;  // per group: index of its sign table among the tables, or -1
    const GroupDesc *gdesc;  // per group: X mask, first Z mask, term range, sign table
    uint32_t n_tab;        // sign tables; entry u of table t at u * n_tab + t
    const void *tab_re;    // sign tables in the precision of coef: real parts
    const void *tab_im;    //   and imaginary parts
    const float *tab_lo_re;  // FP32 sign tables (dual fill)
    const float *tab_lo_im;
    const void *diag;      // double2[N] or float2[N]
    const uint32_t *tile_row0;
    const uint32_t *tile_seg;
    const uint32_t *tile_prob;     // problem of each tile
    const ProbDesc *prob;
    unsigned int *queue;           // work-queue counter (persistent launch)
    int32_t *idx;                  // lookup output
    double2 *vals;                 // lookup output (optional): matrix element C_j(s_i)
    unsigned long long *counts;    // count output
    const uint64_t *g_x;           // full X mask per group
    const void *gy;                // gradient mode: upstream gradient (same dtype as x)
    double *grad;                  // gradient mode: per-term output (accumulated)
    uint32_t *rowcnt;              // row-count mode: section-B hits per row
    uint32_t *rowcnt_a;            // row-count mode: section-A hits per row
    // hit cache passes.  Column pass: columns of the full packed layout (slice s
    // from c_off[s], section A then B), the group of each section-A entry (-1 on
    // padding; level k of slice s at (c_aoff[s] + k) * 64 + lane) and the values
    // of section B (as in the value pass).  Value
    // pass: values straight into the stored layout (section B packed from
    // c_boff[s]; section-A level k by its code c_acode[c_aoff[s] + k]: -1 for one
    // value at c_aval[c_aoff[s] + k], else its index j among the slice's levels
    // with 64 values, at c_vala[c_avoff[s] + 64 j + lane]).
    int32_t *c_col;
    int32_t *c_gid;
    const int64_t *c_off;
    const uint32_t *c_wa;          // section-A entries per row of each slice
    const uint32_t *c_wb;          // section-B entries per row of each slice (the widest row)
    const uint8_t *c_bpos;         // section B: each row's position in its slice (SELL order)
    const uint16_t *c_bw;          // section B: the 8 chunk widths (packets) of each slice
    const int64_t *c_aoff;
    const int64_t *c_boff;
    const int64_t *c_avoff;
    const int32_t *c_acode;
    void *c_val;                   // section-B values (R, or complex R)
    void *c_aval;
    void *c_vala;
    void *c_val_lo;                // (R = double, optional) FP32 values of the same hits
    void *c_aval_lo;
    void *c_vala_lo;
    const void *coef_lo;           // with c_val_lo: FP32 coefficients
    int32_t c_a_only;              // value pass: section A only (the column pass wrote section B's values)
    unsigned long long *coo_count; // coo mode: entries claimed so far
    int64_t *coo_row;              // coo mode: (row, column, element) of every hit, in claim order
    int32_t *coo_col;
    double2 *coo_val;
    int64_t coo_capacity;
    int64_t n_total;
    uint32_t n_tiles;
    uint32_t hq_off;               // matvec: byte offset of the per-wave hit queues in the dynamic LDS
    // alpha lists (matvec): entries per alpha index, first entry of each block of
    // 64 alpha indices, the entries in the matvec's precision (AlEnt<R>), and per
    // class the groups of an alpha class outside the lists ([al_pg_off[c], al_pg_off[c + 1]))
    const uint32_t *al_len;
    const int64_t *al_boff;
    const void *al_ent;
    const uint32_t *al_pg_off;
    const uint32_t *al_pg;
    const AlEnt<float> *al_ent_lo;  // the FP32 entries (a cache fill of both precisions)
    const uint32_t *al_eg;          // per alpha-list entry: its group
    const void *al_ec;              // the entries for real coefficients (AlEntC, precision of coef)
    const SplitDesc *al_split;      // per group (split groups only)
    const void *al_G;               // G tables in the precision of coef (double2 or float2)
    const float2 *al_G_lo;          // FP32 G tables (a cache fill of both precisions)
    // partner lists (product subspaces): per alpha index its partners (alpha
    // pattern, alpha pattern index, a', 0) in the blocked layout of the alpha
    // lists; per beta index its partners (beta pattern index, b') at
    // [al_bpoff[b], al_bpoff[b + 1]); the partner map: the group of each (beta
    // pattern, alpha pattern) pair, or -1
    const uint32_t *al_plen;
    const int64_t *al_pboff;
    const uint4 *al_pent;
    const uint32_t *al_bpoff;
    const uint2 *al_bpent;
    const int32_t *al_pmap;
    const uint64_t *al_ap;          // per alpha pattern (the plan's list): the pattern and its
    const uint64_t *al_pz;          //   reference Z part (the partners' alpha signs)
    // beta lists (product subspaces whose lists cover every group): per beta index
    // its hits of the groups that flip only beta bits, [bl_off[b], bl_off[b + 1]),
    // as (group, b' | AL_SPLIT, value) in the precision of coef (and FP32)
    const uint32_t *bl_off;
    const void *bl_ent;
    const AlEnt<float> *bl_ent_lo;
};

template <typename R> struct C2;
template <> struct C2<double> { using T = double2; };
template <> struct C2<float> { using T = float2; };

template <typename R, bool CPLX> struct VT { using T = R; };
template <typename R> struct VT<R, true> { using T = typename C2<R>::T; };

template <typename R> __device__ inline R vzero(R *) { return R(0); }
__device__ inline double2 vzero(double2 *) { return make_double2(0.0, 0.0); }
__device__ inline float2 vzero(float2 *) { return make_float2(0.f, 0.f); }

// The value of a split beta-list entry at alpha string a: F (the entry's value)
// plus the group's G terms of a (G rows of n values), in pattern order
template <typename R>
__device__ inline typename C2<R>::T split_value(typename C2<R>::T v, const SplitDesc &sd, uint64_t s,
                                                const typename C2<R>::T *G, uint32_t n, uint32_t a) {
#pragma unroll
    for (int j = 0; j < SPLIT_MAX; ++j)
        if (static_cast<uint32_t>(j) < sd.n_p) {
            const typename C2<R>::T gv = G[static_cast<size_t>(sd.grow) + static_cast<size_t>(j) * n + a];
            const bool neg = __popcll(s & sd.zeta[j]) & 1;
            v.x += neg ? -gv.x : gv.x;
            v.y += neg ? -gv.y : gv.y;
        }
    return v;
}

// The value of a split alpha-list entry of beta block b: F (the entry's value)
// plus the group's G terms, Gb = the block's G values, in pattern order
template <typename T> __device__ inline T split_alpha(T v, uint32_t sw, const T *Gb) {
    const uint32_t r = sw & ((1u << SW_ROW_BITS) - 1u), np = (sw >> SW_ROW_BITS) & 3u;
#pragma unroll
    for (int j = 0; j < SPLIT_MAX; ++j)
        if (static_cast<uint32_t>(j) < np) {
            const T gv = Gb[r + j];
            const bool neg = (sw >> (30 + j)) & 1u;
            v.x += neg ? -gv.x : gv.x;
            v.y += neg ? -gv.y : gv.y;
        }
    return v;
}

// acc += c * v  (c complex coefficient sum; v real or complex vector entry)
__device__ inline void fma_acc(double &acc, double2 c, double v) { acc = fma(c.x, v, acc); }
__device__ inline void fma_acc(float &acc, float2 c, float v) { acc = fmaf(c.x, v, acc); }
__device__ inline void fma_acc(double2 &acc, double2 c, double2 v) {
    acc.x = fma(c.x, v.x, acc.x);
    acc.x = fma(-c.y, v.y, acc.x);
    acc.y = fma(c.x, v.y, acc.y);
    acc.y = fma(c.y, v.x, acc.y);
}
__device__ inline void fma_acc(float2 &acc, float2 c, float2 v) {
    acc.x = fmaf(c.x, v.x, acc.x);
    acc.x = fmaf(-c.y, v.y, acc.x);
    acc.y = fmaf(c.x, v.y, acc.y);
    acc.y = fmaf(c.y, v.x, acc.y);
}

// real coefficient (a Hamiltonian whose phased coefficients are all real): acc += c * v
__device__ inline void fma_acc(double &acc, double c, double v) { acc = fma(c, v, acc); }
__device__ inline void fma_acc(float &acc, float c, float v) { acc = fmaf(c, v, acc); }
__device__ inline void fma_acc(double2 &acc, double c, double2 v) {
    acc.x = fma(c, v.x, acc.x);
    acc.y = fma(c, v.y, acc.y);
}
__device__ inline void fma_acc(float2 &acc, float c, float2 v) {
    acc.x = fmaf(c, v.x, acc.x);
    acc.y = fmaf(c, v.y, acc.y);
}

__device__ inline double vadd(double a, double b) { return a + b; }
__device__ inline float vadd(float a, float b) { return a + b; }
__device__ inline double2 vadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline float2 vadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }

// Map the hardware dispatch order (round robin over XCDs) to contiguous tile
// ranges per XCD, so neighbouring tiles share one XCD's L2.
__device__ inline uint32_t xcd_remap(uint32_t b, uint32_t nb) {
    const uint32_t xcd = b % NUM_XCD, local = b / NUM_XCD;
    const uint32_t q = nb / NUM_XCD, r = nb % NUM_XCD;
    return (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

// index of t in sorted a[0..n) or -1
// Hashed segments: the low keys of a segment sit in a two-choice cuckoo table
// of 2^lg buckets of two u32 entries (key << HASH_RANK_BITS | rank within the
// segment; empty = all ones, never a key with L <= 18), so a lookup reads two
// 8-byte buckets that do not depend on each other.
constexpr int HASH_RANK_BITS = 13;     // segments hold at most 8192 rows
constexpr uint32_t HASH_EMPTY = 0xFFFFFFFFu;
constexpr int HASH_MAX_KEY_BITS = 18;
static_assert(HASH_MAX_KEY_BITS + HASH_RANK_BITS < 32, "an empty entry (all ones) never matches a key");
static_assert(HASH_MAX_KEY_BITS <= 24, "keys are 24-bit multiply operands");

// Multiplicative hashes of a key (at most 18 bits) with 24-bit operands: full-rate
// 24-bit multiplies on the GPU, where a 32-bit multiply issues at a quarter of the rate.
__host__ __device__ inline uint32_t hash_b1(uint32_t k, uint32_t lg) {
    return ((k & 0xFFFFFFu) * 0x9E3779u) >> (32u - lg);
}
__host__ __device__ inline uint32_t hash_b2(uint32_t k, uint32_t lg) {
    return ((k & 0xFFFFFFu) * 0x85EBCBu) >> (32u - lg);
}

// look t up in the staged table; on a hit, rank = its position in the segment
// (on a miss rank is meaningless: every caller uses it only for hits).
// entry ^ (t << HASH_RANK_BITS) is the entry's rank when the entry holds t and at
// least 2^HASH_RANK_BITS otherwise (empty entries too: keys have at most 18 bits),
// so the smallest of the four is the answer.
__device__ inline bool hash_find(const uint32_t *tab, uint32_t lg, uint32_t t, uint32_t &rank) {
    const uint2 a = reinterpret_cast<const uint2 *>(tab)[hash_b1(t, lg)];
    const uint2 b = reinterpret_cast<const uint2 *>(tab)[hash_b2(t, lg)];
    const uint32_t tk = t << HASH_RANK_BITS;
    const uint32_t m = min(min(a.x ^ tk, a.y ^ tk), min(b.x ^ tk, b.y ^ tk));
    rank = m;
    return m < (1u << HASH_RANK_BITS);
}

__device__ inline int find_u64(const uint64_t *__restrict__ a, int n, uint64_t t) {
    if (n <= 0) return -1;
    int base = 0, len = n;
    while (len > 1) {
        const int half = len >> 1;
        base = (a[base + half] <= t) ? base + half : base;
        len -= half;
    }
    return (a[base] == t) ? base : -1;
}

// C_g(s) = sum_k coef_k (-1)^popcount(s & z_k), coefficients in precision R
template <typename R>
__device__ inline typename C2<R>::T group_coef_range(const void *coef_v, const KParams &p, uint32_t k0, uint32_t k1,
                                                     uint64_t s) {
    using CC = typename C2<R>::T;
    const CC *coef = static_cast<const CC *>(coef_v);
    CC acc;
    acc.x = R(0);
    acc.y = R(0);
    // the terms' (z, coef) are wave-uniform scalar loads, which return out of
    // order: fetch TERM_PAD terms before using any, so one scalar-memory round
    // trip serves them all (the term arrays carry TERM_PAD entries of padding;
    // the sum keeps its term order)
    for (uint32_t k = k0; k < k1; k += TERM_PAD) {
        uint64_t z[TERM_PAD];
        CC c[TERM_PAD];
#pragma unroll
        for (int u = 0; u < TERM_PAD; ++u) {
            z[u] = ldc(p.z + k + u);
            c[u] = ldc(coef + k + u);
        }
        __builtin_amdgcn_sched_barrier(0);   // keep the loads ahead of their first use
#pragma unroll
        for (int u = 0; u < TERM_PAD; ++u) {
            if (k + u < k1) {
                const bool odd = __popcll(s & z[u]) & 1;
                acc.x += odd ? -c[u].x : c[u].x;
                acc.y += odd ? -c[u].y : c[u].y;
            }
        }
    }
    return acc;
}

template <typename R>
__device__ inline typename C2<R>::T group_coef_from(const void *coef_v, const KParams &p, uint32_t g, uint64_t s) {
    return group_coef_range<R>(coef_v, p, ldc(p.g_toff + g), ldc(p.g_toff + g + 1), s);
}

template <typename R>
__device__ inline typename C2<R>::T group_coef(const KParams &p, uint32_t g, uint64_t s) {
    return group_coef_from<R>(p.coef, p, g, s);
}

// Re C_g(s), the whole sum when the coefficients are real (same order as group_coef_range)
template <typename R>
__device__ inline R group_coef_re_range(const void *coef_v, const KParams &p, uint32_t k0, uint32_t k1, uint64_t s) {
    using CC = typename C2<R>::T;
    const CC *coef = static_cast<const CC *>(coef_v);
    R acc = R(0);
    for (uint32_t k = k0; k < k1; k += TERM_PAD) {
        uint64_t z[TERM_PAD];
        R c[TERM_PAD];
#pragma unroll
        for (int u = 0; u < TERM_PAD; ++u) {
            z[u] = ldc(p.z + k + u);
            c[u] = ldc(coef + k + u).x;
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < TERM_PAD; ++u) {
            if (k + u < k1) {
                const bool odd = __popcll(s & z[u]) & 1;
                acc += odd ? -c[u] : c[u];
            }
        }
    }
    return acc;
}

template <typename R>
__device__ inline R group_coef_re_from(const void *coef_v, const KParams &p, uint32_t g, uint64_t s) {
    return group_coef_re_range<R>(coef_v, p, ldc(p.g_toff + g), ldc(p.g_toff + g + 1), s);
}

// sum over the 64 lanes of a wave (every lane receives the total)
template <typename R>
__device__ inline R wave_sum(R v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Coefficient sums C_g(s) of one group for the lanes with `hit`, evaluated by
// the whole wave: row-parallel (each hitting lane sums the group's terms in
// order) or, when few lanes hit a long group, term-parallel (the 64 lanes split
// the terms of one hitting row -- lane j takes terms j, j + 64, ... -- and a
// butterfly sum joins them; one hitting row after the other).  The choice
// depends only on the group's length and the number of hitting lanes, so every
// pass over the same tiles (matvec, cache fill) evaluates a hit the same way.
__device__ inline bool term_parallel(uint32_t n_terms, uint32_t n_hits) {
    return n_hits * (((n_terms + 63u) >> 6) * 10u + 20u) < n_terms * 8u;
}
// Sign tables.  When every term of a group has z_k = z_0 ^ t_k with t_k inside
// the group's X mask x (JW double excitations: z differs only on the four
// flipped orbitals), C(s) = (-1)^popc(s & z_0) tab[s restricted to x], where
// tab[u] = sum_k c_k (-1)^popc(u & t_k) over the 2^popc(x) patterns u, summed
// in term order.  Negation commutes with round-to-nearest, so this equals the
// term-by-term sum bit for bit.
constexpr int TAB_MAX_BITS = 4;      // tables of at most 16 entries
constexpr int TAB_ENTRIES = 1 << TAB_MAX_BITS;
constexpr int TAB_MIN_TERMS = 3;

// the bits of s at a group's X positions (packed in its descriptor), packed low:
// the sign-table pattern of s
__device__ inline uint32_t tab_index_packed(uint32_t xpos, uint64_t s) {
    const uint32_t nb = xpos >> 24;
    uint32_t idx = 0;
#pragma unroll
    for (int j = 0; j < TAB_MAX_BITS; ++j)
        if (static_cast<uint32_t>(j) < nb) idx |= static_cast<uint32_t>((s >> ((xpos >> (6 * j)) & 63u)) & 1ull) << j;
    return idx;
}

// the entry of s's pattern in group d's sign table: pattern u of table t sits at
// u * n_tab + t (n_tab < 2^24, so a full-rate 24-bit multiply gives it)
constexpr int64_t TAB_MAX_TABLES = int64_t(1) << 24;
__device__ inline uint32_t tab_entry(const GroupDesc &d, uint64_t s, uint32_t n_tab) {
    return (tab_index_packed(d.xpos, s) & 0xFFu) * (n_tab & 0xFFFFFFu) + static_cast<uint32_t>(d.tab);
}

// the rule never picks term-parallel for groups shorter than this (8 K > 30
// at one hit); plans without such groups launch matvec kernels compiled with
// the row-parallel path only (TP = false), which give the same sums
constexpr int TP_MIN_TERMS = 4;

template <typename R, bool RE, bool TP = true>
__device__ inline void group_value_wave(const void *coef_v, const KParams &p, const GroupDesc &d, uint64_t s,
                                        bool hit, R &vx, R &vy) {
    using CC = typename C2<R>::T;
    vx = R(0);
    vy = R(0);
    if constexpr (!TP) {
        if (hit) {
            if constexpr (RE) {
                vx = group_coef_re_range<R>(coef_v, p, d.k0, d.k1, s);
            } else {
                const CC c = group_coef_range<R>(coef_v, p, d.k0, d.k1, s);
                vx = c.x;
                vy = c.y;
            }
        }
        return;
    }
    if (d.tab >= 0) {
        if (hit) {
            const uint32_t u = tab_entry(d, s, p.n_tab);
            const bool neg = __popcll(s & d.z0) & 1;
            const R re = static_cast<const R *>(p.tab_re)[u];
            vx = neg ? -re : re;
            if constexpr (!RE) {
                const R im = static_cast<const R *>(p.tab_im)[u];
                vy = neg ? -im : im;
            }
        }
        return;
    }
    const uint64_t m = __ballot(hit);
    if (m == 0) return;
    const uint32_t k0 = d.k0, k1 = d.k1;
    if (!term_parallel(k1 - k0, static_cast<uint32_t>(__popcll(m)))) {
        if (hit) {
            if constexpr (RE) {
                vx = group_coef_re_range<R>(coef_v, p, k0, k1, s);
            } else {
                const CC c = group_coef_range<R>(coef_v, p, k0, k1, s);
                vx = c.x;
                vy = c.y;
            }
        }
        return;
    }
    const CC *coef = static_cast<const CC *>(coef_v);
    const int lane = threadIdx.x & 63;
    for (uint64_t mm = m; mm != 0; mm &= mm - 1) {
        const int L = __ffsll(static_cast<long long>(mm)) - 1;
        const uint64_t sl = __shfl(s, L, 64);
        R px = R(0), py = R(0);
        for (uint32_t k = k0 + lane; k < k1; k += 64) {
            const CC c = coef[k];
            const bool odd = __popcll(sl & p.z[k]) & 1;
            px += odd ? -c.x : c.x;
            if constexpr (!RE) py += odd ? -c.y : c.y;
        }
        px = wave_sum(px);
        if constexpr (!RE) py = wave_sum(py);
        if (lane == L) {
            vx = px;
            vy = py;
        }
    }
}

template <typename R, bool RE, bool TP = true>
__device__ inline void group_value_wave(const void *coef_v, const KParams &p, uint32_t g, uint64_t s, bool hit,
                                        R &vx, R &vy) {
    if constexpr (TP) {
        group_value_wave<R, RE, TP>(coef_v, p, ldc(p.gdesc + g), s, hit, vx, vy);
    } else {            // the plain evaluator needs only the term range
        vx = R(0);
        vy = R(0);
        if (hit) {
            if constexpr (RE) {
                vx = group_coef_re_from<R>(coef_v, p, g, s);
            } else {
                const typename C2<R>::T c = group_coef_from<R>(coef_v, p, g, s);
                vx = c.x;
                vy = c.y;
            }
        }
    }
}

// Matvec hit lists (one per wave, in the dynamic LDS after the staged target):
// each lane lists its own hits of the groups a lane can evaluate alone -- group
// and rank in the target, in group order -- and evaluates them when a lane's
// list is full or the class ends, one round per list position, adding each to
// its row in group order.  A hit costs one LDS store where it is found; the
// per-group work of the wave does not grow with the number of hitting lanes.
// Entry k of lane l at k * 64 + l (consecutive lanes, consecutive words).
constexpr int HQ_LANE = 8;
template <typename R, bool CPLX, bool REALC> struct HQLayout {
    static constexpr size_t meta = 0;                                  // uint2: group, rank
    static constexpr size_t bytes = 64 * HQ_LANE * 8;                  // per wave
};
constexpr size_t HQ_LDS_MAX = (TILE / 64) * HQLayout<double, true, false>::bytes;   // the largest, per workgroup

// C_g(s) of one (row, group) pair evaluated by one lane, for the groups whose
// sum group_value_wave never splits over the wave: a sign-table group, or one of
// fewer than TP_MIN_TERMS terms (summed in term order, as group_coef_range does,
// with per-lane loads since the lanes of a wave evaluate different groups).
// Equal, bit for bit, to what group_value_wave gives that row.
__device__ inline bool lane_evaluable(const GroupDesc &d) {
    return d.tab >= 0 || d.k1 - d.k0 < static_cast<uint32_t>(TP_MIN_TERMS);
}

template <typename R, bool RE>
__device__ inline void lane_group_value(const KParams &p, const GroupDesc &d, uint64_t s, R &vx, R &vy) {
    using CC = typename C2<R>::T;
    vx = R(0);
    vy = R(0);
    if (d.tab >= 0) {
        const uint32_t u = tab_entry(d, s, p.n_tab);
        const bool neg = __popcll(s & d.z0) & 1;
        const R re = static_cast<const R *>(p.tab_re)[u];
        vx = neg ? -re : re;
        if constexpr (!RE) {
            const R im = static_cast<const R *>(p.tab_im)[u];
            vy = neg ? -im : im;
        }
        return;
    }
    const CC *coef = static_cast<const CC *>(p.coef);
    for (uint32_t k = d.k0; k < d.k1; ++k) {
        const CC c = coef[k];
        const bool odd = __popcll(s & p.z[k]) & 1;
        vx += odd ? -c.x : c.x;
        if constexpr (!RE) vy += odd ? -c.y : c.y;
    }
}

// lane_group_value in FP32 from the FP32 coefficients and tables of a fill of
// both precisions (coef_lo, tab_lo_re, tab_lo_im): what the FP32 matvec gives
template <bool RE>
__device__ inline void lane_group_value_lo(const KParams &p, const GroupDesc &d, uint64_t s, float &vx, float &vy) {
    vx = 0.f;
    vy = 0.f;
    if (d.tab >= 0) {
        const uint32_t u = tab_entry(d, s, p.n_tab);
        const bool neg = __popcll(s & d.z0) & 1;
        const float re = p.tab_lo_re[u];
        vx = neg ? -re : re;
        if constexpr (!RE) {
            const float im = p.tab_lo_im[u];
            vy = neg ? -im : im;
        }
        return;
    }
    const float2 *coef = static_cast<const float2 *>(p.coef_lo);
    for (uint32_t k = d.k0; k < d.k1; ++k) {
        const float2 c = coef[k];
        const bool odd = __popcll(s & p.z[k]) & 1;
        vx += odd ? -c.x : c.x;
        if constexpr (!RE) vy += odd ? -c.y : c.y;
    }
}

// the same in FP64 and FP32 at once (the cache fill of both precisions): each
// precision summed exactly as group_value_wave sums it
template <bool RE>
__device__ inline void group_value_wave_dual(const KParams &p, const GroupDesc &d, uint64_t s, bool hit, double2 &hi,
                                             float2 &lo);

// real and imaginary parts of conj(g) * v for real or complex entries
__device__ inline void conj_mul(double g, double v, double &re, double &im) { re = g * v; im = 0.0; }
__device__ inline void conj_mul(float g, float v, float &re, float &im) { re = g * v; im = 0.f; }
__device__ inline void conj_mul(double2 g, double2 v, double &re, double &im) {
    re = g.x * v.x + g.y * v.y;
    im = g.x * v.y - g.y * v.x;
}
__device__ inline void conj_mul(float2 g, float2 v, float &re, float &im) {
    re = g.x * v.x + g.y * v.y;
    im = g.x * v.y - g.y * v.x;
}

// Re((-i)^ph * (re + i im))
template <typename R>
__device__ inline R re_phase(int ph, R re, R im) {
    return (ph == 0) ? re : (ph == 1) ? im : (ph == 2) ? -re : -im;
}

// Gradient of Re<gy, H x> with respect to the unphased coefficient of every
// term of group g: sum over rows of Re(conj(gy_i) (-i)^{x.z} (-1)^{s_i.z} x[src]).
// Called by every lane of the wave; lanes without a hit contribute 0.
template <typename R>
__device__ inline void grad_group(const KParams &p, uint32_t g, uint64_t s, bool hit, R wre, R wim) {
    if (__ballot(hit) == 0) return;
    const uint64_t xg = ldc(p.g_x + g);
    const uint32_t k1 = ldc(p.g_toff + g + 1);
    for (uint32_t k = ldc(p.g_toff + g); k < k1; ++k) {
        const uint64_t zk = ldc(p.z + k);
        const int ph = __popcll(xg & zk) & 3;
        R val = hit ? re_phase(ph, wre, wim) : R(0);
        if (__popcll(s & zk) & 1) val = -val;
        val = wave_sum(val);
        if ((threadIdx.x & 63) == 0) atomicAdd(&p.grad[k], static_cast<double>(val));
    }
}

// C_g(s) in FP64 and in FP32 (each summed in its own precision, term by term in
// the order of group_coef_from) from one pass over the terms: the fill of a
// cache built in both precisions
template <bool RE_ONLY>
__device__ inline void group_coef_dual(const KParams &p, uint32_t k0, uint32_t k1, uint64_t s, double2 &hi, float2 &lo) {
    const double2 *c64 = static_cast<const double2 *>(p.coef);
    const float2 *c32 = static_cast<const float2 *>(p.coef_lo);
    hi = make_double2(0.0, 0.0);
    lo = make_float2(0.f, 0.f);
    for (uint32_t k = k0; k < k1; k += TERM_PAD) {
        uint64_t z[TERM_PAD];
        double2 ch[TERM_PAD];
        float2 cl[TERM_PAD];
#pragma unroll
        for (int u = 0; u < TERM_PAD; ++u) {
            z[u] = ldc(p.z + k + u);
            ch[u] = ldc(c64 + k + u);
            cl[u] = ldc(c32 + k + u);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < TERM_PAD; ++u) {
            if (k + u < k1) {
                const bool odd = __popcll(s & z[u]) & 1;
                hi.x += odd ? -ch[u].x : ch[u].x;
                lo.x += odd ? -cl[u].x : cl[u].x;
                if constexpr (!RE_ONLY) {
                    hi.y += odd ? -ch[u].y : ch[u].y;
                    lo.y += odd ? -cl[u].y : cl[u].y;
                }
            }
        }
    }
}

// Append (row, col, value) for the lanes with hit set: one atomic per wave for
// the wave's entries, positions in lane order; entries past the capacity are
// counted but not written.
__device__ inline void coo_emit(const KParams &p, bool hit, int64_t row, int64_t col, double2 v) {
    const uint64_t m = __ballot(hit);
    if (m == 0) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll(static_cast<long long>(m)) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(p.coo_count, static_cast<unsigned long long>(__popcll(m)));
    base = __shfl(base, leader, 64);
    if (hit) {
        const int64_t pos = static_cast<int64_t>(base) + __popcll(m & ((1ull << lane) - 1ull));
        if (pos < p.coo_capacity) {
            p.coo_row[pos] = row;
            p.coo_col[pos] = static_cast<int32_t>(col);
            p.coo_val[pos] = v;
        }
    }
}

template <bool RE>
__device__ inline void group_value_wave_dual(const KParams &p, const GroupDesc &d, uint64_t s, bool hit, double2 &hi,
                                             float2 &lo) {
    hi = make_double2(0.0, 0.0);
    lo = make_float2(0.f, 0.f);
    if (d.tab >= 0) {
        if (hit) {
            const uint32_t u = tab_entry(d, s, p.n_tab);
            const bool neg = __popcll(s & d.z0) & 1;
            const double hre = static_cast<const double *>(p.tab_re)[u];
            const float lre = p.tab_lo_re[u];
            hi.x = neg ? -hre : hre;
            lo.x = neg ? -lre : lre;
            if constexpr (!RE) {
                const double him = static_cast<const double *>(p.tab_im)[u];
                const float lim = p.tab_lo_im[u];
                hi.y = neg ? -him : him;
                lo.y = neg ? -lim : lim;
            }
        }
        return;
    }
    const uint64_t m = __ballot(hit);
    if (m == 0) return;
    const uint32_t k0 = d.k0, k1 = d.k1;
    if (!term_parallel(k1 - k0, static_cast<uint32_t>(__popcll(m)))) {
        if (hit) group_coef_dual<RE>(p, k0, k1, s, hi, lo);
        return;
    }
    const double2 *c64 = static_cast<const double2 *>(p.coef);
    const float2 *c32 = static_cast<const float2 *>(p.coef_lo);
    const int lane = threadIdx.x & 63;
    for (uint64_t mm = m; mm != 0; mm &= mm - 1) {
        const int L = __ffsll(static_cast<long long>(mm)) - 1;
        const uint64_t sl = __shfl(s, L, 64);
        double hx = 0.0, hy = 0.0;
        float lx = 0.f, ly = 0.f;
        for (uint32_t k = k0 + lane; k < k1; k += 64) {
            const bool odd = __popcll(sl & p.z[k]) & 1;
            const double2 ch = c64[k];
            const float2 cl = c32[k];
            hx += odd ? -ch.x : ch.x;
            lx += odd ? -cl.x : cl.x;
            if constexpr (!RE) {
                hy += odd ? -ch.y : ch.y;
                ly += odd ? -cl.y : cl.y;
            }
        }
        hx = wave_sum(hx);
        lx = wave_sum(lx);
        if constexpr (!RE) {
            hy = wave_sum(hy);
            ly = wave_sum(ly);
        }
        if (lane == L) {
            hi = make_double2(hx, hy);
            lo = make_float2(lx, ly);
        }
    }
}

// One tile: <= 256 consecutive rows of one segment, one row per thread.
// REALC: every phased coefficient is real (the matvec then sums real parts only).
// TP: the full evaluator (sign tables, term-parallel sums); a plan with neither
// sign tables nor groups of TP_MIN_TERMS or more terms (random Pauli sums: one
// term per X mask) launches the plain one, which gives the same sums.
template <int MODE, typename R, bool CPLX, bool REALC = false, bool TP = true>
__device__ inline void process_tile(const KParams &p, uint32_t tile, uint32_t *s_keys,
                                    const void *__restrict__ xin, void *__restrict__ yout) {
    using V = typename VT<R, CPLX>::T;
    using CC = typename C2<R>::T;
    const int tid = threadIdx.x;
    const ProbDesc pd = ldc(p.prob + ldc(p.tile_prob + tile));
    const uint32_t seg = ldc(p.tile_seg + tile);
    const uint32_t i = ldc(p.tile_row0 + tile) + tid;
    const uint32_t seg0 = ldc(p.seg_off + seg);       // the segment's first row
    const bool active = i < ldc(p.seg_off + seg + 1);
    const uint64_t s = active ? p.keys[i] : 0ull;
    const uint32_t slo = static_cast<uint32_t>(s) & pd.lomask;
    const V *x = static_cast<const V *>(xin);
    // matvec with the full evaluator: the wave's hit queue
    constexpr bool HQ = MODE == MODE_MATVEC && TP;
    using HL = HQLayout<R, CPLX, REALC>;
    const int lane = tid & 63;
    uint2 *hq = nullptr;
    uint32_t my_n = 0;             // listed hits of this lane
    if constexpr (HQ)
        hq = reinterpret_cast<uint2 *>(reinterpret_cast<char *>(s_keys) + p.hq_off +
                                       static_cast<size_t>(tid >> 6) * HL::bytes + HL::meta);

    V acc = vzero(static_cast<V *>(nullptr));       // diagonal + section-B groups
    V acc_a = vzero(static_cast<V *>(nullptr));     // section-A groups
    // alpha lists (the matvec, the hit count and the cache passes; a value pass
    // of section A only has no use for them): this lane's alpha index a (row
    // i = al_rb + a), its entries [0, al_n) and the next one (group al_g, column
    // al_rb + al_col, split word al_sw, value al_v and, in a fill of both
    // precisions, al_l; al_g = UINT32_MAX when none is left)
    constexpr bool AL_FILL = MODE == MODE_FILL || MODE == MODE_FILLCOL;
    constexpr bool AL = HQ || MODE == MODE_ROWCOUNT || AL_FILL;
    const bool al_on = AL && pd.al_nlo != 0 && !(MODE == MODE_FILL && p.c_a_only);
    uint32_t al_rb = 0, al_k = 0, al_n = 0, al_g = UINT32_MAX, al_col = 0, al_sw = 0, al_b = 0;   // al_b: the beta block
    int64_t al_pos = 0;
    uint32_t pl_n = 0;             // partners of this lane's alpha string, the first at pl_pos
    int64_t pl_pos = 0;
    CC al_v{};
    float2 al_l{};
    auto al_load = [&]() {
        if (al_k < al_n) {
            const int64_t o = al_pos + 64 * static_cast<int64_t>(al_k);
            const AlEnt<R> en = static_cast<const AlEnt<R> *>(p.al_ent)[o];
            al_g = p.al_eg[o];
            al_sw = en.sw;
            al_col = en.col;
            al_v = en.v;
            if constexpr (AL_FILL) {
                if (p.c_val_lo) al_l = p.al_ent_lo[o].v;
            }
        } else {
            al_g = UINT32_MAX;
        }
    };
    unsigned long long n_hit = 0, n_probe = 0, n_look = 0, n_hit_a = 0;
    using FV = typename std::conditional<CPLX, CC, R>::type;          // fill: stored value
    using FL = typename std::conditional<CPLX, float2, float>::type;  // fill: FP32 value of a dual fill
    FillBuf<FV, FL> fill_b;
    const int ln = static_cast<int>(i % SLICE);
    // the row's slice: first full-layout entry (column pass), first section-B
    // value and first stored section-A value (value pass), first section-A level
    int64_t fill_off = 0, fill_bo = 0, fill_avo = 0, fill_ao = 0;
    uint32_t fill_wa = 0;
    // section B (SELL): the row's position in its slice and the slice's chunk widths
    int fill_pos = 0;
    uint32_t fill_W[NCH] = {};
    if constexpr (MODE == MODE_FILLCOL) {
        if (active) {
            fill_off = p.c_off[i / SLICE];
            fill_wa = p.c_wa[i / SLICE];
            fill_ao = p.c_aoff[i / SLICE];
            fill_bo = p.c_boff[i / SLICE];
            fill_pos = p.c_bpos[i];
            load_chunk_widths(p.c_bw, i / SLICE, fill_W);
        }
    } else if constexpr (MODE == MODE_FILL) {
        if (active) {
            fill_bo = p.c_boff[i / SLICE];
            fill_avo = p.c_avoff[i / SLICE];
            fill_ao = p.c_aoff[i / SLICE];
            fill_pos = p.c_bpos[i];
            load_chunk_widths(p.c_bw, i / SLICE, fill_W);
        }
    }
    // section A, column pass: entry k of this row stored at once (in product
    // subspaces the lanes of a wave write the same entry index, so the stores
    // coalesce)
    auto fill_put_col = [&](uint32_t k, int32_t c, int32_t g) {
        p.c_col[packed_slot<int32_t>(fill_off, k, ln)] = c;
        p.c_gid[(fill_ao + k) * SLICE + ln] = g;
    };
    // section A, value pass: the level's one value (every lane of a level with
    // one value holds the same value) or the row's value of a stored level
    auto fill_put = [&](uint32_t k, FV v, FL l) {
        const int32_t code = p.c_acode[fill_ao + k];
        const int64_t o = code < 0 ? fill_ao + k : fill_avo + static_cast<int64_t>(code) * SLICE + ln;
        static_cast<FV *>(code < 0 ? p.c_aval : p.c_vala)[o] = v;
        if (p.c_val_lo) static_cast<FL *>(code < 0 ? p.c_aval_lo : p.c_vala_lo)[o] = l;
    };
    // section B: append entry k of this row; a whole packet goes to memory
    // (columns of the full layout and SELL values in the column pass, SELL values
    // in a full value pass)
    auto fill_push = [&](FillBuf<FV, FL> &b, uint32_t k, int32_t c, FV v, FL l) {
#pragma unroll
        for (int u = 0; u + 1 < PACK; ++u) {
            b.c[u] = b.c[u + 1];
            b.v[u] = b.v[u + 1];
            b.l[u] = b.l[u + 1];
        }
        b.c[PACK - 1] = c;
        b.v[PACK - 1] = v;
        b.l[PACK - 1] = l;
        if (k % PACK == PACK - 1) {
            const uint32_t kb = k - (PACK - 1);
            if constexpr (MODE == MODE_FILLCOL) store_packet<int32_t>(p.c_col, fill_off, fill_wa + kb, ln, b.c);
            store_packet_sell<FV>(p.c_val, fill_bo, fill_W, kb / PACK, fill_pos, b.v);
            if (p.c_val_lo) store_packet_sell<FL>(p.c_val_lo, fill_bo, fill_W, kb / PACK, fill_pos, b.l);
        }
    };
    // a section-B value of a group a lane evaluates alone, as the matvecs of the
    // cache's precisions evaluate it (lanes may hold different groups)
    auto lane_fill_value = [&](const GroupDesc &d, FV &v, FL &l) {
        if constexpr (std::is_same<R, double>::value) {
            if (p.c_val_lo) {
                double hx, hy;
                float lx, ly;
                lane_group_value<double, !CPLX>(p, d, s, hx, hy);
                lane_group_value_lo<!CPLX>(p, d, s, lx, ly);
                if constexpr (CPLX) {
                    v = CC{hx, hy};
                    l = make_float2(lx, ly);
                } else {
                    v = hx;
                    l = lx;
                }
                return;
            }
        }
        R vx, vy;
        lane_group_value<R, !CPLX>(p, d, s, vx, vy);
        if constexpr (CPLX) v = CC{vx, vy};
        else v = vx;
    };
    if constexpr (AL) {
        if (al_on && active) {
            al_b = (i - pd.row0) / pd.al_nlo;
            const uint32_t a = (i - pd.row0) - al_b * pd.al_nlo, ag = pd.al_a0 + a;
            al_rb = i - a;
            al_n = p.al_len[ag];
            al_pos = p.al_boff[ag >> 6] + (ag & 63u);
            if (pd.pm_base != UINT32_MAX) {
                pl_n = p.al_plen[ag];
                pl_pos = p.al_pboff[ag >> 6] + (ag & 63u);
            }
            if constexpr (MODE == MODE_ROWCOUNT) {
                n_hit += al_n;         // counted, not visited
                al_n = 0;
            }
            al_load();
        }
    }
    // this lane's listed hits of the groups before gend, in group order: added
    // to the row (the next entry's load in flight with this one's x load), or
    // appended to the row's section B
    auto al_emit = [&](uint32_t gend) {
        if constexpr (AL) {
            while (__ballot(al_g < gend) != 0) {
                if (al_g < gend) {
                    const uint32_t col = al_rb + (al_col & ~AL_SPLIT);
                    V xv{};
                    if constexpr (MODE == MODE_MATVEC) xv = x[col];
                    CC v = al_v;
                    float2 l = al_l;
                    if (al_col & AL_SPLIT) {
                        const size_t gb = pd.al_g0 + static_cast<size_t>(al_b) * pd.al_gnr;
                        v = split_alpha(v, al_sw, static_cast<const CC *>(p.al_G) + gb);
                        if constexpr (AL_FILL) {
                            if (p.c_val_lo) l = split_alpha(l, al_sw, p.al_G_lo + gb);
                        }
                    }
                    ++al_k;
                    if constexpr (MODE == MODE_MATVEC) {
                        al_load();
                        if constexpr (REALC) fma_acc(acc, v.x, xv);
                        else fma_acc(acc, v, xv);
                    } else if constexpr (AL_FILL) {
                        al_load();
                        if constexpr (CPLX) fill_push(fill_b, static_cast<uint32_t>(n_hit), static_cast<int32_t>(col), v, l);
                        else fill_push(fill_b, static_cast<uint32_t>(n_hit), static_cast<int32_t>(col), v.x, l.x);
                        n_hit += 1;
                    }
                }
            }
        }
    };
    V gv = vzero(static_cast<V *>(nullptr));
    if constexpr (MODE == MODE_GRAD) {
        if (active) gv = static_cast<const V *>(p.gy)[i];
        if (pd.has_diag) {
            R wre = R(0), wim = R(0);
            if (active) conj_mul(gv, x[i], wre, wim);
            grad_group<R>(p, pd.grp_base, s, active, wre, wim);
        }
    }
    if (pd.has_diag && active && MODE != MODE_GRAD) {
        if constexpr (MODE == MODE_MATVEC) {
            const CC d = static_cast<const CC *>(p.diag)[i];
            fma_acc(acc, d, x[i]);
        } else if constexpr (MODE == MODE_LOOKUP) {
            p.idx[static_cast<size_t>(pd.grp_base) * p.n_total + i] = static_cast<int32_t>(i);
            if (p.vals) p.vals[static_cast<size_t>(pd.grp_base) * p.n_total + i] = static_cast<const double2 *>(p.diag)[i];
        } else if constexpr (MODE == MODE_COUNT) {
            n_hit += 1;
        }
    }
    if constexpr (MODE == MODE_COO) {
        if (pd.has_diag)
            coo_emit(p, active, i, i, active ? static_cast<const double2 *>(p.diag)[i] : make_double2(0.0, 0.0));
    }

    // the classes whose target segment exists for this tile's segment, in class
    // order, with that segment (listed once per plan by k_seg_targets)
    // (a fully listed problem searches no class: its groups come from the lists)
    const bool full_on = AL && pd.full != 0;
    const uint32_t e1 = full_on ? 0u : ldc(p.st_off + seg + 1);
    for (uint32_t e = full_on ? 0u : ldc(p.st_off + seg); e < e1; ++e) {
        {
            const TargetDesc td = ldc(p.st_list + e);
            if constexpr (MODE == MODE_COUNT) n_look += (tid == 0);
            const uint32_t ta = td.ta, tn = td.tn;
            const bool hashed = p.h_off != nullptr;          // the same for the whole plan
            const uint32_t lg = td.lg;
            // an alpha class: its listed groups come from the alpha lists, its
            // other groups [pq0, pq1) of al_pg are searched (the target is staged
            // only for them)
            bool al_cls = false;
            uint32_t pq0 = 0, pq1 = 0;
            if constexpr (AL) {
                if (al_on && (ldc(p.cls_lane + td.cls) & 2u) != 0) {
                    al_cls = true;
                    pq0 = ldc(p.al_pg_off + td.cls);
                    pq1 = ldc(p.al_pg_off + td.cls + 1);
                }
            }
            // a section-A group of a product (it flips high bits only; the class's
            // first group) hits every row, at the row's own rank in the target
            // segment: no search.  A class of that group alone stages nothing.
            const bool a_direct = pd.prod && ldc(p.g_xlo + td.g0) == 0 &&
                                  (ldc(p.g_x + td.g0) & section_mask(pd.a_bits)) == 0;
            // a mapped class: its groups that flip alpha bits come from the partner
            // loop after the classes; only a first group of high bits is taken here
            bool mapped = false;
            if constexpr (AL) {
                mapped = al_on && pd.pm_base != UINT32_MAX && (ldc(p.cls_lane + td.cls) & 4u) != 0;
            }
            const bool staged = (!al_cls || pq0 < pq1) && !(a_direct && td.g1 == td.g0 + 1) && !mapped;   // uniform
            if (staged) {
                if (hashed) {
                    // 16-byte pieces: a table has 2 << lg >= 4 words and starts at a multiple of 4
                    const uint32_t ho = td.ho, ns4 = 1u << (lg - 1);
                    const uint4 *src = reinterpret_cast<const uint4 *>(p.h_tab + ho);
                    for (uint32_t k = tid; k < ns4; k += TILE) reinterpret_cast<uint4 *>(s_keys)[k] = src[k];
                } else {
                    for (uint32_t k = tid; k < tn; k += TILE) s_keys[k] = p.lo[ta + k];
                }
                __syncthreads();
            }
            // position of t in the target segment and whether it is there
            auto search = [&](uint32_t t, uint32_t &base) -> bool {
                if (hashed) return hash_find(s_keys, lg, t, base);
                base = 0;
                for (uint32_t len = tn; len > 1;) {
                    const uint32_t half = len >> 1;
                    base = (s_keys[base + half] <= t) ? base + half : base;
                    len -= half;
                }
                return s_keys[base] == t;
            };
            const uint32_t g0 = td.g0, g1 = td.g1;
            // a hit's value (and, for the matvec, its x entry) first, then its
            // use: the full-evaluator matvec and value pass take four groups'
            // values before using them, so their loads are in flight together;
            // values are used in group order
            struct HitVal {
                R vx, vy;
                double2 vh;
                float2 vl;
                V xv;
            };
            auto value = [&](const GroupDesc &d, uint32_t base, bool hit) {
                HitVal h{};
                if constexpr (MODE == MODE_MATVEC) {
                    group_value_wave<R, REALC, TP>(p.coef, p, d, s, hit, h.vx, h.vy);
                    if (hit) h.xv = x[ta + base];
                } else if constexpr (MODE == MODE_FILL || MODE == MODE_FILLCOL) {
                    if (p.c_val_lo) group_value_wave_dual<!CPLX>(p, d, s, hit, h.vh, h.vl);   // both precisions, R = double
                    else group_value_wave<R, !CPLX>(p.coef, p, d, s, hit, h.vx, h.vy);
                }
                return h;
            };
            auto consume = [&](uint32_t g, uint32_t base, bool hit, bool sec_a, const HitVal &h) {
                if constexpr (MODE == MODE_MATVEC) {
                    if constexpr (REALC) {
                        if (hit) fma_acc(sec_a ? acc_a : acc, h.vx, h.xv);
                    } else {
                        CC c;
                        c.x = h.vx;
                        c.y = h.vy;
                        if (hit) fma_acc(sec_a ? acc_a : acc, c, h.xv);
                    }
                } else if constexpr (MODE == MODE_FILL || MODE == MODE_FILLCOL) {
                    if (hit) {
                        FV v;
                        FL l;
                        if constexpr (CPLX) {
                            v = p.c_val_lo ? CC{static_cast<R>(h.vh.x), static_cast<R>(h.vh.y)} : CC{h.vx, h.vy};
                            l = h.vl;
                        } else {
                            v = p.c_val_lo ? static_cast<R>(h.vh.x) : h.vx;
                            l = h.vl.x;
                        }
                        if (sec_a) {                   // value pass only
                            fill_put(static_cast<uint32_t>(n_hit_a), v, l);
                            n_hit_a += 1;
                        } else {
                            fill_push(fill_b, static_cast<uint32_t>(n_hit), static_cast<int32_t>(ta + base), v, l);
                            n_hit += 1;
                        }
                    }
                }
            };
            // each lane evaluates its listed hits in order and adds them to its row;
            // empties the lists
            auto hq_flush = [&]() {
                if constexpr (HQ) {
                    for (uint32_t k = 0; __ballot(k < my_n) != 0; ++k) {
                        if (k < my_n) {
                            const uint2 mt = hq[k * 64 + static_cast<uint32_t>(lane)];
                            const V xv = x[ta + mt.y];
                            R vx, vy;
                            lane_group_value<R, REALC>(p, p.gdesc[mt.x], s, vx, vy);
                            if constexpr (REALC) {
                                fma_acc(acc, vx, xv);
                            } else {
                                CC c;
                                c.x = vx;
                                c.y = vy;
                                fma_acc(acc, c, xv);
                            }
                        }
                    }
                    my_n = 0;
                }
            };
            // list this group's hit (a group the lanes evaluate alone)
            auto hq_push = [&](uint32_t g, uint32_t base, bool hit) {
                if constexpr (HQ) {
                    if (hit) {
                        hq[my_n * 64 + static_cast<uint32_t>(lane)] = make_uint2(g, base);
                        ++my_n;
                    }
                    if (__ballot(my_n == HQ_LANE) != 0) hq_flush();
                }
            };
            auto handle = [&](uint32_t g, uint32_t base, bool found, bool sec_a) {
                const bool hit = active && found;
                if constexpr (MODE == MODE_MATVEC) {
                    R vx, vy;
                    group_value_wave<R, REALC, TP>(p.coef, p, g, s, hit, vx, vy);
                    if constexpr (REALC) {
                        if (hit) fma_acc(sec_a ? acc_a : acc, vx, x[ta + base]);
                    } else {
                        CC c;
                        c.x = vx;
                        c.y = vy;
                        if (hit) fma_acc(sec_a ? acc_a : acc, c, x[ta + base]);
                    }
                } else if constexpr (MODE == MODE_FILL) {
                    consume(g, base, hit, sec_a, value(ldc(p.gdesc + g), base, hit));
                } else if constexpr (MODE == MODE_FILLCOL) {
                    if (sec_a) {                       // columns and groups; values in the value pass
                        if (hit) {
                            fill_put_col(static_cast<uint32_t>(n_hit_a), static_cast<int32_t>(ta + base),
                                         static_cast<int32_t>(g));
                            n_hit_a += 1;
                        }
                    } else {
                        consume(g, base, hit, false, value(ldc(p.gdesc + g), base, hit));
                    }
                } else if constexpr (MODE == MODE_LOOKUP) {
                    if (hit) {
                        const size_t o = static_cast<size_t>(g) * p.n_total + i;
                        p.idx[o] = static_cast<int32_t>(ta + base);
                        if (p.vals) p.vals[o] = group_coef<double>(p, g, s);
                    }
                } else if constexpr (MODE == MODE_GRAD) {
                    R wre = R(0), wim = R(0);
                    if (hit) conj_mul(gv, x[ta + base], wre, wim);
                    grad_group<R>(p, g, s, hit, wre, wim);
                } else if constexpr (MODE == MODE_ROWCOUNT) {
                    if (sec_a) n_hit_a += hit;
                    else n_hit += hit;
                } else if constexpr (MODE == MODE_COO) {
                    double vx, vy;
                    group_value_wave<double, false>(p.coef, p, g, s, hit, vx, vy);
                    coo_emit(p, hit, i, ta + base, make_double2(vx, vy));
                } else {
                    n_hit += hit;
                    n_probe += active;
                }
            };
            uint32_t g = g0;
            // groups are sorted by X mask, so a class's group without low flips is
            // its first one; it is section A when it also flips no bit below the
            // problem's section split (its bits [L, a_bits) are zero), else section B
            if (!al_cls && g < g1 && ldc(p.g_xlo + g) == 0 && (!mapped || a_direct)) {
                const bool sec_a = (ldc(p.g_x + g) & section_mask(pd.a_bits)) == 0;
                uint32_t base = i - seg0;
                const bool f = a_direct || search(slo, base);
                if (!(MODE == MODE_FILL && p.c_a_only && !sec_a))   // section B was filled by the column pass
                    handle(g, base, f, sec_a);
                ++g;
            }
            if constexpr (MODE == MODE_FILL) {
                if (p.c_a_only) g = g1;       // section B was filled by the column pass
            }
            // four independent searches in lock step: the dependent LDS loads of
            // one search overlap with those of the other three
            auto search4 = [&](uint32_t t0, uint32_t t1, uint32_t t2, uint32_t t3, uint32_t &b0, uint32_t &b1,
                               uint32_t &b2, uint32_t &b3, bool &f0, bool &f1, bool &f2, bool &f3) {
                b0 = b1 = b2 = b3 = 0;
                if (hashed) {
                    f0 = hash_find(s_keys, lg, t0, b0);
                    f1 = hash_find(s_keys, lg, t1, b1);
                    f2 = hash_find(s_keys, lg, t2, b2);
                    f3 = hash_find(s_keys, lg, t3, b3);
                } else {
                    for (uint32_t len = tn; len > 1;) {
                        const uint32_t half = len >> 1;
                        b0 = (s_keys[b0 + half] <= t0) ? b0 + half : b0;
                        b1 = (s_keys[b1 + half] <= t1) ? b1 + half : b1;
                        b2 = (s_keys[b2 + half] <= t2) ? b2 + half : b2;
                        b3 = (s_keys[b3 + half] <= t3) ? b3 + half : b3;
                        len -= half;
                    }
                    f0 = s_keys[b0] == t0;
                    f1 = s_keys[b1] == t1;
                    f2 = s_keys[b2] == t2;
                    f3 = s_keys[b3] == t3;
                }
            };
            if (mapped) g = g1;
            // an alpha class: its groups outside the lists searched in group order,
            // this lane's listed hits taken before each and at the end.  The passes
            // that append to section B search one group at a time (the emitter and
            // its packet stores inlined once); the matvec four at a time, its hit
            // queue left empty.
            if constexpr (AL && !HQ) {
                if (al_cls) {
                    for (uint32_t q = pq0;; ++q) {
                        const uint32_t gq = q < pq1 ? ldc(p.al_pg + q) : g1;
                        al_emit(gq);
                        if (q >= pq1) break;
                        uint32_t b;
                        const bool f = search(slo ^ ldc(p.g_xlo + gq), b);
                        handle(gq, b, f, false);
                    }
                    g = g1;
                }
            }
            if constexpr (HQ) {
                if (al_cls) {
                    auto other = [&](const GroupDesc &d, uint32_t gq, uint32_t b, bool f) {
                        al_emit(gq);
                        const bool h = active && f;
                        if (!lane_evaluable(d)) {
                            consume(gq, b, h, false, value(d, b, h));
                        } else if (h) {
                            R vx, vy;
                            lane_group_value<R, REALC>(p, d, s, vx, vy);
                            const V xv = x[ta + b];
                            if constexpr (REALC) {
                                fma_acc(acc, vx, xv);
                            } else {
                                CC c;
                                c.x = vx;
                                c.y = vy;
                                fma_acc(acc, c, xv);
                            }
                        }
                    };
                    uint32_t q = pq0;
                    for (; q + 4 <= pq1; q += 4) {
                        const uint32_t gq0 = ldc(p.al_pg + q), gq1 = ldc(p.al_pg + q + 1),
                                       gq2 = ldc(p.al_pg + q + 2), gq3 = ldc(p.al_pg + q + 3);
                        const GroupDesc d0 = ldc(p.gdesc + gq0), d1 = ldc(p.gdesc + gq1), d2 = ldc(p.gdesc + gq2),
                                        d3 = ldc(p.gdesc + gq3);
                        uint32_t b0, b1, b2, b3;
                        bool f0, f1, f2, f3;
                        search4(slo ^ (static_cast<uint32_t>(d0.x) & pd.lomask), slo ^ (static_cast<uint32_t>(d1.x) & pd.lomask),
                                slo ^ (static_cast<uint32_t>(d2.x) & pd.lomask), slo ^ (static_cast<uint32_t>(d3.x) & pd.lomask),
                                b0, b1, b2, b3, f0, f1, f2, f3);
                        other(d0, gq0, b0, f0);
                        other(d1, gq1, b1, f1);
                        other(d2, gq2, b2, f2);
                        other(d3, gq3, b3, f3);
                    }
                    for (; q < pq1; ++q) {
                        const uint32_t gq = ldc(p.al_pg + q);
                        const GroupDesc d = ldc(p.gdesc + gq);
                        uint32_t b;
                        const bool f = search(slo ^ (static_cast<uint32_t>(d.x) & pd.lomask), b);
                        other(d, gq, b, f);
                    }
                    al_emit(g1);
                    g = g1;
                }
            }
            // the matvec's loop for a class whose groups the lanes all evaluate
            // alone: the searches need only the low X masks, every hit is queued
            if constexpr (HQ) {
                if ((ldc(p.cls_lane + td.cls) & 1u) != 0) {
                    for (; g + 4 <= g1; g += 4) {
                        uint32_t b0, b1, b2, b3;
                        bool f0, f1, f2, f3;
                        search4(slo ^ ldc(p.g_xlo + g), slo ^ ldc(p.g_xlo + g + 1), slo ^ ldc(p.g_xlo + g + 2),
                                slo ^ ldc(p.g_xlo + g + 3), b0, b1, b2, b3, f0, f1, f2, f3);
                        hq_push(g, b0, active && f0);
                        hq_push(g + 1, b1, active && f1);
                        hq_push(g + 2, b2, active && f2);
                        hq_push(g + 3, b3, active && f3);
                    }
                }
            }
            for (; g + 4 <= g1; g += 4) {
                // the full evaluator fetches the four groups' descriptors at once,
                // ahead of the searches (their X masks give the targets)
                constexpr bool ILP = (MODE == MODE_MATVEC || MODE == MODE_FILL || MODE == MODE_FILLCOL) && TP;
                GroupDesc d0{}, d1{}, d2{}, d3{};
                uint32_t t0, t1, t2, t3;
                if constexpr (ILP) {
                    d0 = ldc(p.gdesc + g);
                    d1 = ldc(p.gdesc + g + 1);
                    d2 = ldc(p.gdesc + g + 2);
                    d3 = ldc(p.gdesc + g + 3);
                    t0 = slo ^ (static_cast<uint32_t>(d0.x) & pd.lomask);
                    t1 = slo ^ (static_cast<uint32_t>(d1.x) & pd.lomask);
                    t2 = slo ^ (static_cast<uint32_t>(d2.x) & pd.lomask);
                    t3 = slo ^ (static_cast<uint32_t>(d3.x) & pd.lomask);
                } else {
                    t0 = slo ^ ldc(p.g_xlo + g);
                    t1 = slo ^ ldc(p.g_xlo + g + 1);
                    t2 = slo ^ ldc(p.g_xlo + g + 2);
                    t3 = slo ^ ldc(p.g_xlo + g + 3);
                }
                uint32_t b0, b1, b2, b3;
                bool f0, f1, f2, f3;
                search4(t0, t1, t2, t3, b0, b1, b2, b3, f0, f1, f2, f3);
                if constexpr (HQ) {
                    // lane-evaluable groups into the queue; a group whose sum the wave
                    // splits (long, no sign table) after the queue is flushed, in
                    // group order
                    const bool h0 = active && f0, h1 = active && f1, h2 = active && f2, h3 = active && f3;
                    auto one = [&](const GroupDesc &d, uint32_t gg, uint32_t b, bool h) {
                        if (lane_evaluable(d)) {
                            hq_push(gg, b, h);
                        } else {
                            hq_flush();
                            consume(gg, b, h, false, value(d, b, h));
                        }
                    };
                    one(d0, g, b0, h0);
                    one(d1, g + 1, b1, h1);
                    one(d2, g + 2, b2, h2);
                    one(d3, g + 3, b3, h3);
                } else if constexpr (ILP) {   // (measured: slower for single-term groups)
                    const bool h0 = active && f0, h1 = active && f1, h2 = active && f2, h3 = active && f3;
                    const HitVal v0 = value(d0, b0, h0), v1 = value(d1, b1, h1);
                    const HitVal v2 = value(d2, b2, h2), v3 = value(d3, b3, h3);
                    consume(g, b0, h0, false, v0);
                    consume(g + 1, b1, h1, false, v1);
                    consume(g + 2, b2, h2, false, v2);
                    consume(g + 3, b3, h3, false, v3);
                } else {
                    handle(g, b0, f0, false);
                    handle(g + 1, b1, f1, false);
                    handle(g + 2, b2, f2, false);
                    handle(g + 3, b3, f3, false);
                }
            }
            hq_flush();               // the queued hits precede the remaining groups
            for (; g < g1; ++g) {
                uint32_t base;
                const bool f = search(slo ^ ldc(p.g_xlo + g), base);
                handle(g, base, f, false);
            }
            if (staged) __syncthreads();   // before the next target overwrites s_keys
        }
    }

    // the mapped classes' groups that flip alpha bits: for each partner of the
    // tile's beta string (the same for every row of a tile; beta patterns in
    // order) and each partner of the row's alpha string (alpha patterns in order),
    // the map's group hits at (b', a')
    if constexpr (AL) {
        if (full_on) al_emit(UINT32_MAX);          // every alpha-list entry (no class took them)
        if (al_on && pd.pm_base != UINT32_MAX) {
            const uint32_t jb = (ldc(p.tile_row0 + tile) - pd.row0) / pd.al_nlo;
            const uint32_t q1 = ldc(p.al_bpoff + pd.bp_b0 + jb + 1);
            for (uint32_t q = ldc(p.al_bpoff + pd.bp_b0 + jb); q < q1; ++q) {
                const uint2 bq = ldc(p.al_bpent + q);          // beta pattern index, b'
                const int32_t *mrow = p.al_pmap + pd.pm_base + static_cast<size_t>(bq.x) * pd.pm_na;
                const uint32_t tb = pd.row0 + bq.y * pd.al_nlo;
                for (uint32_t k = 0; __ballot(k < pl_n) != 0; ++k) {
                    if (k < pl_n) {
                        const uint4 pe = p.al_pent[pl_pos + 64 * static_cast<int64_t>(k)];
                        const int32_t gg = mrow[pe.y];
                        if (gg >= 0) {
                            const uint32_t col = tb + pe.z;
                            const GroupDesc d = p.gdesc[gg];
                            if constexpr (MODE == MODE_MATVEC) {
                                const V xv = x[col];
                                R vx, vy;
                                lane_group_value<R, REALC>(p, d, s, vx, vy);
                                if constexpr (REALC) {
                                    fma_acc(acc, vx, xv);
                                } else {
                                    CC c;
                                    c.x = vx;
                                    c.y = vy;
                                    fma_acc(acc, c, xv);
                                }
                            } else if constexpr (MODE == MODE_ROWCOUNT) {
                                n_hit += 1;
                            } else if constexpr (AL_FILL) {
                                FV fv;
                                FL fl{};
                                lane_fill_value(d, fv, fl);
                                fill_push(fill_b, static_cast<uint32_t>(n_hit), static_cast<int32_t>(col), fv, fl);
                                n_hit += 1;
                            }
                        }
                    }
                }
            }
        }
    }

    // a fully listed problem's groups of high bits only: the tile's beta string's
    // list (the same for every row), each hit at (b', the row's alpha index)
    if constexpr (AL) {
        if (full_on) {
            const uint32_t jb = (ldc(p.tile_row0 + tile) - pd.row0) / pd.al_nlo;
            const uint32_t a_l = i - pd.row0 - jb * pd.al_nlo;      // (active rows)
            const uint32_t q0 = ldc(p.bl_off + pd.bl_b0 + jb), q1 = ldc(p.bl_off + pd.bl_b0 + jb + 1);
            if constexpr (MODE == MODE_ROWCOUNT) {
                if (active) n_hit_a += q1 - q0;
            } else {
                for (uint32_t q = q0; q < q1; ++q) {
                    const AlEnt<R> en = ldc(static_cast<const AlEnt<R> *>(p.bl_ent) + q);
                    const uint32_t col = pd.row0 + (en.col & ~AL_SPLIT) * pd.al_nlo + a_l;
                    CC v = en.v;
                    float2 l{};
                    if constexpr (AL_FILL) {
                        if (p.c_val_lo) l = ldc(p.bl_ent_lo + q).v;
                    }
                    if (en.col & AL_SPLIT) {
                        const SplitDesc sd = ldc(p.al_split + en.g);
                        if (active) {
                            v = split_value<R>(v, sd, s, static_cast<const CC *>(p.al_G), pd.al_nlo, a_l);
                            if constexpr (AL_FILL) {
                                if (p.c_val_lo) l = split_value<float>(l, sd, s, p.al_G_lo, pd.al_nlo, a_l);
                            }
                        }
                    }
                    if (active) {
                        if constexpr (MODE == MODE_MATVEC) {
                            if constexpr (REALC) fma_acc(acc_a, v.x, x[col]);
                            else fma_acc(acc_a, v, x[col]);
                        } else if constexpr (MODE == MODE_FILLCOL) {
                            fill_put_col(static_cast<uint32_t>(n_hit_a), static_cast<int32_t>(col), static_cast<int32_t>(en.g));
                            n_hit_a += 1;
                        } else if constexpr (MODE == MODE_FILL) {
                            if constexpr (CPLX) fill_put(static_cast<uint32_t>(n_hit_a), v, l);
                            else fill_put(static_cast<uint32_t>(n_hit_a), v.x, l.x);
                            n_hit_a += 1;
                        }
                    }
                }
            }
        }
    }

    if constexpr (MODE == MODE_MATVEC) {
        if (active) static_cast<V *>(yout)[i] = vadd(acc, acc_a);
    } else if constexpr (MODE == MODE_ROWCOUNT) {
        if (active) {
            p.rowcnt[i] = static_cast<uint32_t>(n_hit);
            p.rowcnt_a[i] = static_cast<uint32_t>(n_hit_a);
        }
    } else if constexpr (MODE == MODE_FILL || MODE == MODE_FILLCOL) {
        if (active) {
            // pad to the slice's section-A width and to the row's chunk width in
            // section B (multiples of PACK): the last packets go out; padding has
            // the row's own column, group -1 and value 0
            const uint32_t wa = p.c_wa[i / SLICE], wb = fill_W[fill_pos / CH] * PACK;
            const FV v0 = vzero(static_cast<FV *>(nullptr));
            const FL l0 = vzero(static_cast<FL *>(nullptr));
            for (uint32_t k = static_cast<uint32_t>(n_hit_a); k < wa; ++k) {
                if constexpr (MODE == MODE_FILLCOL) fill_put_col(k, static_cast<int32_t>(i), -1);
                else fill_put(k, v0, l0);
            }
            if (!(MODE == MODE_FILL && p.c_a_only))
                for (uint32_t k = static_cast<uint32_t>(n_hit); k < wb; ++k)
                    fill_push(fill_b, k, static_cast<int32_t>(i), v0, l0);
        }
    } else if constexpr (MODE == MODE_COUNT) {
        // wave reduction, one atomic per wave and counter (the class lookups are
        // counted per thread like the rest: no block-leader branch at the tail)
        for (int off = 32; off > 0; off >>= 1) {
            n_hit += __shfl_down(n_hit, off, 64);
            n_probe += __shfl_down(n_probe, off, 64);
            n_look += __shfl_down(n_look, off, 64);
        }
        if ((tid & 63) == 0) {
            atomicAdd(&p.counts[0], n_hit);
            atomicAdd(&p.counts[1], n_probe);
            atomicAdd(&p.counts[2], n_look);
        }
    }
}

// The matvec of a fully listed product problem (list_mode 2): one workgroup of
// PROD_WG threads per chunk of a beta block, a thread per row.  x of the block
// and its G values (split alpha-list entries) are staged in LDS first: every
// alpha-list gather stays inside the row's block, and scattered 8-byte reads from
// LDS cost a few cycles where the texture path takes a cycle per cache line.  Per row the sums of the tile kernel's fully listed path,
// in its order (diagonal, alpha-list entries, partner hits; beta-list entries;
// their sum) and with the same values, so the results are the same bits (the cache
// passes run the tile kernel).
constexpr int PROD_LDS_BYTES = 64 * 1024;

// (WG threads: 1024 for beta blocks of 2048 rows or more, whose staging the 16 waves
// share; 256 for smaller ones, so that small problems still spread over every CU)
// (eight waves per SIMD, two 1024-thread workgroups per CU, with 64 VGPRs; the FP64
// complex-vector variants would spill there and ask for six)
template <typename R, bool CPLX> constexpr int PROD_WAVES = (sizeof(R) == 8 && CPLX) ? 6 : 8;

template <typename R, bool CPLX, bool REALC, int PROD_WG>
__global__ __launch_bounds__(PROD_WG) __attribute__((amdgpu_waves_per_eu(PROD_WAVES<R, CPLX>))) void k_product(
    const KParams p, uint32_t prob, uint32_t chunks,
                                                     const void *__restrict__ xin, void *__restrict__ yout) {
    using V = typename VT<R, CPLX>::T;
    using CC = typename C2<R>::T;
    extern __shared__ int4 prod_lds[];
    V *xs = reinterpret_cast<V *>(prod_lds);
    const ProbDesc pd = ldc(p.prob + prob);
    // the beta block's G values of the split alpha-list entries after the block's x
    CC *gs = reinterpret_cast<CC *>(prod_lds + (static_cast<size_t>(pd.al_nlo) * sizeof(V) + 15) / 16);
    // chunk-major: the workgroups of one XCD take one chunk of alpha strings over
    // consecutive beta blocks, so that chunk's list entries stay in its L2
    const uint32_t w = xcd_remap(blockIdx.x, gridDim.x), nhi = gridDim.x / chunks;
    const uint32_t b = w % nhi, a = (w / nhi) * PROD_WG + threadIdx.x;
    const V *x = static_cast<const V *>(xin);
    const uint32_t blk = pd.row0 + b * pd.al_nlo;          // the beta block's first row
    for (uint32_t k = threadIdx.x; k < pd.al_nlo; k += PROD_WG) xs[k] = x[blk + k];
    {
        const CC *G = static_cast<const CC *>(p.al_G) + pd.al_g0 + static_cast<size_t>(b) * pd.al_gnr;
        for (uint32_t k = threadIdx.x; k < pd.al_gnr; k += PROD_WG) gs[k] = G[k];
    }
    __syncthreads();
    // threads past the block's last row take part in the stagings and barriers only
    const bool act = a < pd.al_nlo;
    const uint32_t i = blk + (act ? a : 0u);
    const uint64_t s = p.keys[i];
    V acc = vzero(static_cast<V *>(nullptr)), acc_a = vzero(static_cast<V *>(nullptr));
    auto add = [&](V &to, const CC &v, const V &xv) {
        if constexpr (REALC) fma_acc(to, v.x, xv);
        else fma_acc(to, v, xv);
    };
    if (act && pd.has_diag) fma_acc(acc, static_cast<const CC *>(p.diag)[i], xs[a]);
    const uint32_t ag = pd.al_a0 + (act ? a : 0u);
    if (act && REALC) {   // alpha lists (real coefficients: the one-word entries), as below
        constexpr int AL_BATCH = 8;
        const AlEntC<R> *al = static_cast<const AlEntC<R> *>(p.al_ec) + p.al_boff[ag >> 6] + (ag & 63u);
        const uint32_t n = p.al_len[ag];
        auto value = [&](const AlEntC<R> &en) {
            R v = en.v;
            const uint32_t np = (en.w >> CE_BITS) & 3u, r = en.w >> (CE_BITS + 4);
#pragma unroll
            for (int j = 0; j < SPLIT_MAX; ++j)    // split_alpha's real part
                if (static_cast<uint32_t>(j) < np) {
                    const R gv = gs[r + j].x;
                    v += ((en.w >> (CE_BITS + 2 + j)) & 1u) ? -gv : gv;
                }
            return v;
        };
        uint32_t k = 0;
        for (; k + AL_BATCH <= n; k += AL_BATCH) {
            AlEntC<R> en[AL_BATCH];
#pragma unroll
            for (int u = 0; u < AL_BATCH; ++u) en[u] = al[64 * static_cast<int64_t>(k + u)];
#pragma unroll
            for (int u = 0; u < AL_BATCH; ++u) fma_acc(acc, value(en[u]), xs[en[u].w & ((1u << CE_BITS) - 1u)]);
        }
        for (; k < n; ++k) {
            const AlEntC<R> en = al[64 * static_cast<int64_t>(k)];
            fma_acc(acc, value(en), xs[en.w & ((1u << CE_BITS) - 1u)]);
        }
    } else if (act) {   // alpha lists, AL_BATCH entries in flight, then their multiply-adds in order
        constexpr int AL_BATCH = sizeof(R) == 4 ? 8 : 4;
        const AlEnt<R> *al = static_cast<const AlEnt<R> *>(p.al_ent) + p.al_boff[ag >> 6] + (ag & 63u);
        const uint32_t n = p.al_len[ag];
        uint32_t k = 0;
        for (; k + AL_BATCH <= n; k += AL_BATCH) {
            AlEnt<R> en[AL_BATCH];
#pragma unroll
            for (int u = 0; u < AL_BATCH; ++u) en[u] = al[64 * static_cast<int64_t>(k + u)];
#pragma unroll
            for (int u = 0; u < AL_BATCH; ++u) {
                CC v = en[u].v;
                if (en[u].col & AL_SPLIT) v = split_alpha(v, en[u].sw, gs);
                add(acc, v, xs[en[u].col & ~AL_SPLIT]);
            }
        }
        for (; k < n; ++k) {
            const AlEnt<R> en = al[64 * static_cast<int64_t>(k)];
            CC v = en.v;
            if (en.col & AL_SPLIT) v = split_alpha(v, en.sw, gs);
            add(acc, v, xs[en.col & ~AL_SPLIT]);
        }
    }
    if (pd.pm_base != UINT32_MAX) {   // partners through the map
        // per beta partner: its block of x staged in LDS (the rows' partner hits
        // gather from all over it) and, per alpha pattern u, the map's group for
        // the beta partner's pattern: hk[u] = -1 (none), -2 (values in hv), or
        // the group (evaluated per lane).  A group with a sign table whose first
        // term's alpha part is the pattern's reference pz_u but for bits zz at the
        // pattern's X bits has C(s) = (-1)^popc(s & z0) tab[u(s)], u(s) = the alpha
        // bits ua | the beta bits of the block's string, and (-1)^popc(s & z0) =
        // (-1)^popc(s_beta & z0) (-1)^popc(ua & zz) (-1)^popc(s_alpha & pz_u):
        // hv[u][ua] holds the first two signs times tab[u(s)], the row's partner
        // entry gives ua and the third sign -- the same value as the row's own
        // evaluation, bit for bit (negation is exact).  Then the row's hits in
        // alpha-pattern order, its first PK partners held in registers
        constexpr int PK = 4;
        const uint32_t hs = 1u << pd.al_hsl;
        CC *hv = reinterpret_cast<CC *>(prod_lds + (static_cast<size_t>(pd.al_nlo) * sizeof(V) + 15) / 16 +
                                        (static_cast<size_t>(pd.al_gnr) * sizeof(CC) + 15) / 16);
        int32_t *hk = reinterpret_cast<int32_t *>(hv + static_cast<size_t>(pd.pm_na) * hs);
        const uint64_t amask = section_mask(pd.a_bits), sb = p.keys[blk] & ~amask;
        const uint32_t pn = act ? p.al_plen[ag] : 0u;
        const uint4 *pl = p.al_pent + p.al_pboff[ag >> 6] + (ag & 63u);
        uint32_t pm[PK], pa[PK], pw[PK];
#pragma unroll
        for (int u = 0; u < PK; ++u)
            if (static_cast<uint32_t>(u) < pn) {
                const uint4 pe = pl[64 * u];
                pm[u] = pe.y;
                pa[u] = pe.z;
                pw[u] = pe.w;
            }
        auto hit = [&](uint32_t m, uint32_t col, uint32_t w) {
            const int32_t kind = hk[m];
            if (kind == -2) {
                CC v = hv[m * hs + (w & 15u)];
                if (w & 16u) {
                    v.x = -v.x;
                    v.y = -v.y;
                }
                add(acc, v, xs[col]);
            } else if (kind >= 0) {
                const GroupDesc d = p.gdesc[kind];
                CC c;
                lane_group_value<R, REALC>(p, d, s, c.x, c.y);
                add(acc, c, xs[col]);
            }
        };
        const uint32_t q1 = ldc(p.al_bpoff + pd.bp_b0 + b + 1);
        for (uint32_t q = ldc(p.al_bpoff + pd.bp_b0 + b); q < q1; ++q) {
            const uint2 bq = ldc(p.al_bpent + q);
            const int32_t *mrow = p.al_pmap + pd.pm_base + static_cast<size_t>(bq.x) * pd.pm_na;
            const uint32_t tb = pd.row0 + bq.y * pd.al_nlo;
            __syncthreads();                                 // the previous block is read
            for (uint32_t k = threadIdx.x; k < pd.al_nlo; k += PROD_WG) xs[k] = x[tb + k];
            for (uint32_t u = threadIdx.x; u < pd.pm_na; u += PROD_WG) {
                const int32_t g = mrow[u];
                int32_t kind = g;
                if (g >= 0) {
                    const GroupDesc d = p.gdesc[g];
                    const uint64_t xa = p.al_ap[pd.al_pat0 + u], zz = (d.z0 & amask) ^ p.al_pz[pd.al_pat0 + u];
                    const uint32_t na = static_cast<uint32_t>(__popcll(xa));
                    if (d.tab >= 0 && na <= pd.al_hsl && (zz & ~xa) == 0) {
                        uint32_t mz = 0;                     // zz over the table index's alpha bits
                        uint64_t t = xa;
                        for (uint32_t j = 0; j < na; ++j, t &= t - 1ull)
                            mz |= static_cast<uint32_t>((zz >> __builtin_ctzll(t)) & 1ull) << j;
                        const uint32_t ub = tab_index_packed(d.xpos, sb);
                        const bool nb = __popcll(sb & d.z0) & 1;
                        for (uint32_t ua = 0; ua < (1u << na); ++ua) {
                            const uint32_t e = ((ub | ua) & 0xFFu) * (p.n_tab & 0xFFFFFFu) + static_cast<uint32_t>(d.tab);
                            const R re = static_cast<const R *>(p.tab_re)[e];
                            R im = R(0);
                            if constexpr (!REALC) im = static_cast<const R *>(p.tab_im)[e];
                            const bool neg = nb ^ (__popc(ua & mz) & 1);
                            CC v;
                            v.x = neg ? -re : re;
                            v.y = neg ? -im : im;
                            hv[u * hs + ua] = v;
                        }
                        kind = -2;
                    }
                }
                hk[u] = kind;
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < PK; ++u)
                if (static_cast<uint32_t>(u) < pn) hit(pm[u], pa[u], pw[u]);
            for (uint32_t k = PK; k < pn; ++k) {
                const uint4 pe = pl[64 * static_cast<int64_t>(k)];
                hit(pe.y, pe.z, pe.w);
            }
        }
    }
    if (!act) return;                                      // (after the last barrier)
    {   // beta lists: BL_BATCH entries (scalar loads) and their x loads in flight, then
        // their multiply-adds in order
        constexpr int BL_BATCH = 8;
        const AlEnt<R> *bl = static_cast<const AlEnt<R> *>(p.bl_ent);
        const uint32_t q1 = ldc(p.bl_off + pd.bl_b0 + b + 1);
        uint32_t q = ldc(p.bl_off + pd.bl_b0 + b);
        for (; q + BL_BATCH <= q1; q += BL_BATCH) {
            AlEnt<R> en[BL_BATCH];
            V xv[BL_BATCH];
#pragma unroll
            for (int u = 0; u < BL_BATCH; ++u) en[u] = ldc(bl + q + u);
#pragma unroll
            for (int u = 0; u < BL_BATCH; ++u) xv[u] = x[pd.row0 + (en[u].col & ~AL_SPLIT) * pd.al_nlo + a];
#pragma unroll
            for (int u = 0; u < BL_BATCH; ++u) {
                CC v = en[u].v;
                if (en[u].col & AL_SPLIT)
                    v = split_value<R>(v, ldc(p.al_split + en[u].g), s, static_cast<const CC *>(p.al_G), pd.al_nlo, a);
                add(acc_a, v, xv[u]);
            }
        }
        for (; q < q1; ++q) {
            const AlEnt<R> en = ldc(bl + q);
            CC v = en.v;
            if (en.col & AL_SPLIT)
                v = split_value<R>(v, ldc(p.al_split + en.g), s, static_cast<const CC *>(p.al_G), pd.al_nlo, a);
            add(acc_a, v, x[pd.row0 + (en.col & ~AL_SPLIT) * pd.al_nlo + a]);
        }
    }
    static_cast<V *>(yout)[i] = vadd(acc, acc_a);
}

// Occupancy hints.  The cache passes with real values are register-bound (both
// precisions' values, packet buffers): six waves per SIMD (a few registers
// spill) measured 6% faster than five; complex values would spill too much.
// The matvec is bound by the latency of its loads: asking for eight waves per
// SIMD (the compiler then keeps within 64 VGPRs and 96 SGPRs) measured 3-6%
// faster with the full evaluator (seven: the same) and 11% with the plain one;
// the FP64 complex-coefficient evaluator would spill at eight and runs at six
// (6% faster than at eight).
template <int MODE, typename R, bool CPLX, bool REALC>
constexpr int TILE_WAVES = ((MODE == MODE_FILL || MODE == MODE_FILLCOL) && !CPLX) ? 6
                           : MODE != MODE_MATVEC                                  ? 1
                           : (sizeof(R) == 8 && CPLX && !REALC)                   ? 6   // would spill at 8
                                                                                  : 8;

// One workgroup per tile; tiles of one XCD are contiguous.
template <int MODE, typename R, bool CPLX, bool REALC = false, bool TP = true>
__global__ __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(TILE_WAVES<MODE, R, CPLX, REALC>))) void k_tiles(const KParams p, const void *__restrict__ xin,
                                                void *__restrict__ yout) {
    extern __shared__ uint32_t s_keys[];      // staged target segment (max_segment keys)
    process_tile<MODE, R, CPLX, REALC, TP>(p, xcd_remap(blockIdx.x, gridDim.x), s_keys, xin, yout);
}

// Persistent workgroups pull tiles from a queue (tiles are stored largest
// first); every workgroup leaves when the queue is exhausted.
template <int MODE, typename R, bool CPLX, bool REALC = false, bool TP = true>
__global__ __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(TILE_WAVES<MODE, R, CPLX, REALC>))) void k_tiles_queue(const KParams p, const void *__restrict__ xin,
                                                      void *__restrict__ yout) {
    extern __shared__ uint32_t s_keys[];
    __shared__ uint32_t s_tile;
    // a workgroup can claim at most n_tiles + 1 times (the last claim is past the end)
    for (uint32_t claim = 0; claim <= p.n_tiles; ++claim) {
        if (threadIdx.x == 0) s_tile = atomicAdd(p.queue, 1u);
        __syncthreads();
        // readfirstlane makes the claimed tile a scalar: the loop exit below is then
        // a uniform branch, so the compiler keeps every lane of every wave on the
        // same path back to the claim (a divergent-looking exit let it park lane 0
        // outside the loop while the other lanes re-read the previous claim)
        const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);
        __syncthreads();              // every thread has read s_tile before the next claim
        if (tile >= p.n_tiles) break;
        process_tile<MODE, R, CPLX, REALC, TP>(p, tile, s_keys, xin, yout);
        __syncthreads();              // the whole workgroup leaves the tile before the next claim
    }
}


// y[begin, end) = value; grid-stride, so any length fits a capped grid
template <typename V>
__global__ void k_fill(V *__restrict__ y, int64_t begin, int64_t end, V value) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = begin + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < end; i += stride)
        y[i] = value;
}

template <typename V>
hipError_t fill_async(V *y, int64_t begin, int64_t end, V value, hipStream_t st) {
    if (end <= begin) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((end - begin + 255) / 256, int64_t(1) << 16);
    hipLaunchKernelGGL(k_fill<V>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, y, begin, end, value);
    return hipGetLastError();
}

// (class, target segment) pairs of one problem's segments: one workgroup per
// segment m, the problem's classes c in chunks of 256, kept when the segment
// seg_hi[m] ^ cls_hi[c] exists, in class order.  list == NULL: count only
// (count[m - seg0]); else write from st_off[m] on, with the target's rows and
// hash table (h_off == NULL: none) and the class's groups (skipping the group
// skip_grp, the problem's x == 0 group, if it is the class's first).
__global__ __launch_bounds__(256) void k_seg_targets(const uint64_t *__restrict__ seg_hi,
                                                     const uint64_t *__restrict__ cls_hi, uint32_t seg0,
                                                     uint32_t n_segs, uint32_t cls0, uint32_t n_classes,
                                                     uint32_t *__restrict__ count, const uint32_t *__restrict__ st_off,
                                                     const uint32_t *__restrict__ seg_off,
                                                     const uint32_t *__restrict__ h_off,
                                                     const uint32_t *__restrict__ h_log,
                                                     const uint32_t *__restrict__ cls_goff, uint32_t skip_grp,
                                                     TargetDesc *__restrict__ list) {
    __shared__ uint32_t s_wave[TILE / 64];
    const uint32_t m = seg0 + blockIdx.x;
    const uint64_t shi = seg_hi[m];
    const int tid = static_cast<int>(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    uint32_t written = 0;
    for (uint32_t cb = 0; cb < n_classes; cb += TILE) {
        const uint32_t c = cb + static_cast<uint32_t>(tid);
        int t = -1;
        if (c < n_classes) t = find_u64(seg_hi + seg0, static_cast<int>(n_segs), shi ^ cls_hi[cls0 + c]);
        const uint64_t bal = __ballot(t >= 0);
        if (lane == 0) s_wave[wave] = static_cast<uint32_t>(__popcll(bal));
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int w = 0; w < TILE / 64; ++w) {
            before += (w < wave) ? s_wave[w] : 0u;
            total += s_wave[w];
        }
        if (t >= 0 && list) {
            const uint32_t below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(bal >> 32),
                                                             __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(bal), 0u));
            const uint32_t ts = seg0 + static_cast<uint32_t>(t), cls = cls0 + c;
            TargetDesc d;
            d.ta = seg_off[ts];
            d.tn = seg_off[ts + 1] - d.ta;
            d.ho = h_off ? h_off[ts] : 0u;
            d.lg = h_off ? h_log[ts] : 0u;
            d.g0 = cls_goff[cls];
            d.g1 = cls_goff[cls + 1];
            if (d.g0 == skip_grp) d.g0 += 1;
            d.cls = cls;
            d.ts = ts;
            list[st_off[m] + written + before + below] = d;
        }
        written += total;
        __syncthreads();              // s_wave is rewritten by the next chunk
    }
    if (tid == 0 && count) count[blockIdx.x] = written;
}

// sign tables of the groups that have one (one thread per group), in FP64 and
// FP32, each summed in term order as the term-by-term evaluation sums
__global__ void k_group_tables(const int32_t *__restrict__ g_tab, const uint64_t *__restrict__ g_x,
                               const uint32_t *__restrict__ g_toff, const uint64_t *__restrict__ z,
                               const double2 *__restrict__ c64, const float2 *__restrict__ c32, int32_t n_groups,
                               uint32_t n_tab, double *__restrict__ t64re, double *__restrict__ t64im,
                               float *__restrict__ t32re, float *__restrict__ t32im) {
    const int32_t g = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= n_groups) return;
    const int32_t t = g_tab[g];
    if (t < 0) return;
    const uint64_t x = g_x[g];
    const uint32_t k0 = g_toff[g], k1 = g_toff[g + 1];
    const uint64_t z0 = z[k0];
    const uint32_t n_pat = 1u << __popcll(x);
    for (uint32_t u = 0; u < n_pat; ++u) {
        uint64_t sb = 0;              // pattern u deposited on the bits of x
        int j = 0;
        for (uint64_t xm = x; xm != 0; xm &= xm - 1, ++j)
            if ((u >> j) & 1u) sb |= xm & (~xm + 1ull);
        double2 a = make_double2(0.0, 0.0);
        float2 f = make_float2(0.f, 0.f);
        for (uint32_t k = k0; k < k1; ++k) {
            const bool odd = __popcll(sb & (z[k] ^ z0)) & 1;
            const double2 c = c64[k];
            const float2 d = c32[k];
            a.x += odd ? -c.x : c.x;
            a.y += odd ? -c.y : c.y;
            f.x += odd ? -d.x : d.x;
            f.y += odd ? -d.y : d.y;
        }
        // pattern-major: the few patterns a subspace uses are shared by all
        // tables, so their entries of neighbouring groups share cache lines
        const size_t o = static_cast<size_t>(u) * n_tab + static_cast<uint32_t>(t);
        t64re[o] = a.x;
        t64im[o] = a.y;
        t32re[o] = f.x;
        t32im[o] = f.y;
    }
}

// F of a split group at s: its terms whose Z masks lie inside keep (the group's
// own side), in term order
template <typename R>
__device__ inline typename C2<R>::T split_f(const KParams &p, const GroupDesc &d, uint64_t s, uint64_t keep) {
    using CC = typename C2<R>::T;
    const CC *coef = static_cast<const CC *>(p.coef);
    CC v;
    v.x = R(0);
    v.y = R(0);
    for (uint32_t k = d.k0; k < d.k1; ++k) {
        const uint64_t z = p.z[k];
        if ((z & ~keep) != 0) continue;
        const CC c = coef[k];
        const bool odd = __popcll(s & z) & 1;
        v.x += odd ? -c.x : c.x;
        v.y += odd ? -c.y : c.y;
    }
    return v;
}

// G tables of one problem's split groups: row r (group rg[r], pattern rz[r]) at
// entry e = the sum over the group's terms with Z bits in smask (the other side's
// bits) and pattern rz[r] outside smask of c_k (-1)^popc(s_e & z_k), s_e = keys[e *
// stride] & smask; at G[roff[r] + e * estride], FP64 and FP32 (each in term
// order).  Alpha groups: smask = the beta bits, entries per beta block (stride
// n_lo), a block's rows contiguous (estride = the problem's rows); beta groups:
// smask = the alpha bits, entries per alpha string (stride 1), a row's entries
// contiguous (estride 1).
__global__ __launch_bounds__(256) void k_split_g(const KParams p64, const KParams p32, const uint64_t *__restrict__ keys,
                                                 uint32_t stride, uint32_t n, uint32_t estride, uint64_t smask,
                                                 const uint32_t *__restrict__ rg, const uint64_t *__restrict__ rz,
                                                 const uint32_t *__restrict__ roff, double2 *__restrict__ G64,
                                                 float2 *__restrict__ G32) {
    const uint32_t r = blockIdx.y, e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const uint64_t se = keys[static_cast<size_t>(e) * stride] & smask, zeta = rz[r];
    const GroupDesc d = p64.gdesc[rg[r]];
    const double2 *c64 = static_cast<const double2 *>(p64.coef);
    const float2 *c32 = static_cast<const float2 *>(p32.coef);
    double2 h = make_double2(0.0, 0.0);
    float2 l = make_float2(0.f, 0.f);
    for (uint32_t k = d.k0; k < d.k1; ++k) {
        const uint64_t z = p64.z[k];
        if ((z & smask) == 0 || (z & ~smask) != zeta) continue;
        const bool odd = __popcll(se & z) & 1;
        const double2 ch = c64[k];
        const float2 cl = c32[k];
        h.x += odd ? -ch.x : ch.x;
        h.y += odd ? -ch.y : ch.y;
        l.x += odd ? -cl.x : cl.x;
        l.y += odd ? -cl.y : cl.y;
    }
    const size_t o = roff[r] + static_cast<size_t>(e) * estride;
    G64[o] = h;
    G32[o] = l;
}

// Beta lists of one problem: workgroup b takes beta index b (state keys[b * nlo])
// and the problem's beta groups lg[0..nlg) (they flip only beta bits) in order:
// group g hits when beta_b ^ x_g is a beta string, at index b'.  Count pass (off
// NULL): cnt[b] = the hits.  Fill pass: hit k at off[b] + k as (g, b' | AL_SPLIT
// for a split group, value at keys[b * nlo]: lane_group_value, or F of a split
// group -- its terms without alpha Z bits), FP64 and FP32.
__global__ __launch_bounds__(256) void k_beta_lists(const KParams p64, const KParams p32, const uint64_t *__restrict__ keys,
                                                    uint32_t nlo, uint32_t nhi, uint64_t amask,
                                                    const uint32_t *__restrict__ lg, uint32_t nlg,
                                                    uint32_t *__restrict__ cnt, const uint32_t *__restrict__ off,
                                                    AlEnt<float> *__restrict__ e32, AlEnt<double> *__restrict__ e64) {
    __shared__ uint32_t wn[4];
    const uint32_t b = blockIdx.x, tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint64_t s = keys[static_cast<size_t>(b) * nlo], sb = s & ~amask;
    uint32_t base = 0;
    for (uint32_t j0 = 0; j0 < nlg; j0 += 256) {
        const uint32_t j = j0 + tid;
        bool f = false;
        uint32_t g = 0, col = 0;
        if (j < nlg) {
            g = lg[j];
            const uint64_t t = sb ^ p64.g_x[g];
            uint32_t lo = 0, n = nhi;          // lower bound of t among the beta strings
            while (n > 0) {
                const uint32_t half = n >> 1;
                if ((keys[static_cast<size_t>(lo + half) * nlo] & ~amask) < t) {
                    lo += half + 1;
                    n -= half + 1;
                } else {
                    n = half;
                }
            }
            f = lo < nhi && (keys[static_cast<size_t>(lo) * nlo] & ~amask) == t;
            col = lo;
        }
        const uint64_t mk = __ballot(f);
        if (lane == 0) wn[w] = static_cast<uint32_t>(__popcll(mk));
        __syncthreads();
        uint32_t before = base;
        for (uint32_t u = 0; u < w; ++u) before += wn[u];
        const uint32_t total = wn[0] + wn[1] + wn[2] + wn[3];
        if (off && f) {
            const uint32_t pos = off[b] + before + static_cast<uint32_t>(__popcll(mk & ((1ull << lane) - 1ull)));
            const GroupDesc gd = p64.gdesc[g];
            AlEnt<double> eh{};
            AlEnt<float> el{};
            eh.g = el.g = g;
            if (p64.al_split[g].split) {
                eh.col = el.col = col | AL_SPLIT;
                eh.v = split_f<double>(p64, gd, s, ~amask);
                el.v = split_f<float>(p32, gd, s, ~amask);
            } else {
                double hx, hy;
                float lx, ly;
                lane_group_value<double, false>(p64, gd, s, hx, hy);
                lane_group_value<float, false>(p32, gd, s, lx, ly);
                eh.col = el.col = col;
                eh.v = make_double2(hx, hy);
                el.v = make_float2(lx, ly);
            }
            e64[pos] = eh;
            e32[pos] = el;
        }
        base += total;
        __syncthreads();          // wn is rewritten by the next chunk
    }
    if (!off && tid == 0) cnt[b] = base;
}

// Partner lists of one problem: workgroup a takes alpha index a and the sorted
// alpha parts pat[0..npat) of the groups that flip both alpha and beta bits;
// pattern m is a partner when alpha_a ^ pat[m] is an alpha string, at index a'.
// Count pass (boff NULL): len[a0 + a] = the partners.  Fill pass: partner k at
// boff[(a0 + a) >> 6] + 64 k + ((a0 + a) & 63) as (pat[m], m, a', w), ascending m,
// w = alpha_a's bits at pat[m]'s positions (the first 4, ascending: the alpha bits
// of a sign-table index) | popc(alpha_a & pz[m]) mod 2 in bit 4.
__global__ __launch_bounds__(256) void k_alpha_partners(const uint64_t *__restrict__ keys, uint32_t nlo, uint64_t amask,
                                                        const uint64_t *__restrict__ pat,
                                                        const uint64_t *__restrict__ pz, uint32_t npat, uint32_t a0,
                                                        uint32_t *__restrict__ len, const int64_t *__restrict__ boff,
                                                        uint4 *__restrict__ ent) {
    __shared__ uint32_t wn[4];
    const uint32_t a = blockIdx.x, tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint64_t sa = keys[a] & amask;
    const uint32_t ag = a0 + a;
    uint32_t base = 0;
    for (uint32_t m0 = 0; m0 < npat; m0 += 256) {
        const uint32_t m = m0 + tid;
        bool f = false;
        uint32_t col = 0;
        if (m < npat) {
            const uint64_t t = sa ^ pat[m];
            uint32_t lo = 0, n = nlo;          // lower bound of t
            while (n > 0) {
                const uint32_t half = n >> 1;
                if ((keys[lo + half] & amask) < t) {
                    lo += half + 1;
                    n -= half + 1;
                } else {
                    n = half;
                }
            }
            f = lo < nlo && (keys[lo] & amask) == t;
            col = lo;
        }
        const uint64_t mk = __ballot(f);
        if (lane == 0) wn[w] = static_cast<uint32_t>(__popcll(mk));
        __syncthreads();
        uint32_t before = base;
        for (uint32_t u = 0; u < w; ++u) before += wn[u];
        const uint32_t total = wn[0] + wn[1] + wn[2] + wn[3];
        if (boff && f) {
            const uint32_t k = before + static_cast<uint32_t>(__popcll(mk & ((1ull << lane) - 1ull)));
            uint64_t t = pat[m];
            uint32_t w4 = 0;
            for (int j = 0; j < TAB_MAX_BITS && t; ++j, t &= t - 1ull)
                w4 |= static_cast<uint32_t>((sa >> __builtin_ctzll(t)) & 1ull) << j;
            w4 |= (static_cast<uint32_t>(__popcll(sa & pz[m])) & 1u) << 4;
            ent[boff[ag >> 6] + 64 * static_cast<int64_t>(k) + (ag & 63u)] =
                make_uint4(static_cast<uint32_t>(pat[m]), m, col, w4);
        }
        base += total;
        __syncthreads();          // wn is rewritten by the next chunk
    }
    if (!boff && tid == 0) len[ag] = base;
}

// Beta partners of one problem: workgroup b takes beta index b (state keys[b *
// nlo]) and the sorted beta parts bpat[0..nbpat) of the mapped groups; pattern m
// is a partner when its beta string ^ bpat[m] is a beta string, at index b'.
// Count pass (off NULL): cnt[b] = the partners.  Fill pass: partner k at
// off[b] + k as (m, b'), ascending m.
__global__ __launch_bounds__(256) void k_beta_partners(const uint64_t *__restrict__ keys, uint32_t nlo, uint32_t nhi,
                                                       uint64_t amask, const uint64_t *__restrict__ bpat, uint32_t nbpat,
                                                       uint32_t *__restrict__ cnt, const uint32_t *__restrict__ off,
                                                       uint2 *__restrict__ ent) {
    __shared__ uint32_t wn[4];
    const uint32_t b = blockIdx.x, tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint64_t sb = keys[static_cast<size_t>(b) * nlo] & ~amask;
    uint32_t base = 0;
    for (uint32_t m0 = 0; m0 < nbpat; m0 += 256) {
        const uint32_t m = m0 + tid;
        bool f = false;
        uint32_t col = 0;
        if (m < nbpat) {
            const uint64_t t = sb ^ bpat[m];
            uint32_t lo = 0, n = nhi;          // lower bound of t among the beta strings
            while (n > 0) {
                const uint32_t half = n >> 1;
                if ((keys[static_cast<size_t>(lo + half) * nlo] & ~amask) < t) {
                    lo += half + 1;
                    n -= half + 1;
                } else {
                    n = half;
                }
            }
            f = lo < nhi && (keys[static_cast<size_t>(lo) * nlo] & ~amask) == t;
            col = lo;
        }
        const uint64_t mk = __ballot(f);
        if (lane == 0) wn[w] = static_cast<uint32_t>(__popcll(mk));
        __syncthreads();
        uint32_t before = base;
        for (uint32_t u = 0; u < w; ++u) before += wn[u];
        const uint32_t total = wn[0] + wn[1] + wn[2] + wn[3];
        if (off && f) ent[off[b] + before + static_cast<uint32_t>(__popcll(mk & ((1ull << lane) - 1ull)))] = make_uint2(m, col);
        base += total;
        __syncthreads();          // wn is rewritten by the next chunk
    }
    if (!off && tid == 0) cnt[b] = base;
}

// Alpha lists of one problem: workgroup a takes alpha index a (row a of the
// first beta block, keys[a]) and the problem's listed groups lg[0..nlg) in
// order: group g hits when alpha_a ^ (x_g & amask) is an alpha string, at its
// index a' among keys[0..nlo) (ascending low parts).  Count pass (boff NULL):
// len[a0 + a] = the hits.  Fill pass: entry k of the hits at
// boff[(a0 + a) >> 6] + 64 k + ((a0 + a) & 63), with C_g(s) as the matvec's lanes
// evaluate it (lane_group_value) in FP64 and in FP32 -- or, for a split group, F
// and the split word (its first G row, pattern count and alpha_a's signs) --
// and g at the same position of eg.
__global__ __launch_bounds__(256) void k_alpha_lists(const KParams p64, const KParams p32,
                                                     const uint64_t *__restrict__ keys, uint32_t nlo, uint64_t amask,
                                                     const uint32_t *__restrict__ lg, uint32_t nlg, uint32_t a0,
                                                     uint32_t *__restrict__ len, const int64_t *__restrict__ boff,
                                                     AlEnt<float> *__restrict__ e32, AlEnt<double> *__restrict__ e64,
                                                     uint32_t *__restrict__ eg) {
    __shared__ uint32_t wn[4];
    const uint32_t a = blockIdx.x, tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint64_t s = keys[a], sa = s & amask;
    const uint32_t ag = a0 + a;
    uint32_t base = 0;
    for (uint32_t j0 = 0; j0 < nlg; j0 += 256) {
        const uint32_t j = j0 + tid;
        bool f = false;
        uint32_t g = 0, col = 0;
        if (j < nlg) {
            g = lg[j];
            const uint64_t t = sa ^ (p64.g_x[g] & amask);
            uint32_t lo = 0, n = nlo;          // lower bound of t
            while (n > 0) {
                const uint32_t half = n >> 1;
                if ((keys[lo + half] & amask) < t) {
                    lo += half + 1;
                    n -= half + 1;
                } else {
                    n = half;
                }
            }
            f = lo < nlo && (keys[lo] & amask) == t;
            col = lo;
        }
        const uint64_t m = __ballot(f);
        if (lane == 0) wn[w] = static_cast<uint32_t>(__popcll(m));
        __syncthreads();
        uint32_t before = base;
        for (uint32_t u = 0; u < w; ++u) before += wn[u];
        const uint32_t total = wn[0] + wn[1] + wn[2] + wn[3];
        if (boff && f) {
            const uint32_t k = before + static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
            const int64_t pos = boff[ag >> 6] + 64 * static_cast<int64_t>(k) + (ag & 63u);
            const GroupDesc gd = p64.gdesc[g];
            AlEnt<double> eh{};
            AlEnt<float> el{};
            const SplitDesc sd = p64.al_split[g];
            if (sd.split) {
                uint32_t sw = 0;
                if (sd.n_p) {
                    sw = sd.grow | sd.n_p << SW_ROW_BITS;
                    for (uint32_t j = 0; j < sd.n_p; ++j) sw |= (__popcll(s & sd.zeta[j]) & 1u) << (30 + j);
                }
                eh.sw = el.sw = sw;
                eh.col = el.col = sd.n_p ? col | AL_SPLIT : col;
                eh.v = split_f<double>(p64, gd, s, amask);
                el.v = split_f<float>(p32, gd, s, amask);
            } else {
                double hx, hy;
                float lx, ly;
                lane_group_value<double, false>(p64, gd, s, hx, hy);
                lane_group_value<float, false>(p32, gd, s, lx, ly);
                eh.col = el.col = col;
                eh.v = make_double2(hx, hy);
                el.v = make_float2(lx, ly);
            }
            e64[pos] = eh;
            e32[pos] = el;
            eg[pos] = g;
        }
        base += total;
        __syncthreads();          // wn is rewritten by the next chunk
    }
    if (!boff && tid == 0) len[ag] = base;
}

// The real-coefficient form of alpha-list slots [0, n) (padding slots included:
// never read)
__global__ void k_compact_alpha(const AlEnt<float> *__restrict__ e32, const AlEnt<double> *__restrict__ e64,
                                AlEntC<float> *__restrict__ c32, AlEntC<double> *__restrict__ c64, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const AlEnt<float> el = e32[i];
    const AlEnt<double> eh = e64[i];
    uint32_t w = (el.col & ~AL_SPLIT) & ((1u << CE_BITS) - 1u);
    if (el.col & AL_SPLIT)
        w |= ((el.sw >> SW_ROW_BITS) & 3u) << CE_BITS | (el.sw >> 30) << (CE_BITS + 2) |
             (el.sw & ((1u << SW_ROW_BITS) - 1u)) << (CE_BITS + 4);
    c32[i] = AlEntC<float>{w, el.v.x};
    c64[i] = AlEntC<double>{w, 0u, eh.v.x};
}

__global__ void k_narrow(const double2 *__restrict__ in, float2 *__restrict__ out, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float2(static_cast<float>(in[i].x), static_cast<float>(in[i].y));
}

// diag[i] = C_0(s_i) for group 0 (x == 0), in double and float copies
__global__ void k_diag(const uint64_t *__restrict__ keys, int64_t n, const uint64_t *__restrict__ z,
                       const double2 *__restrict__ coef, uint32_t k0, uint32_t k1,
                       double2 *__restrict__ d64, float2 *__restrict__ d32) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t s = keys[i];
    double re = 0.0, im = 0.0;
    for (uint32_t k = k0; k < k1; ++k) {
        const bool odd = __popcll(s & z[k]) & 1;
        re += odd ? -coef[k].x : coef[k].x;
        im += odd ? -coef[k].y : coef[k].y;
    }
    d64[i] = make_double2(re, im);
    d32[i] = make_float2(static_cast<float>(re), static_cast<float>(im));
}

// ---------------------------------------------------------------- small eigenpairs
struct cd {
    double r, i;
};
__host__ __device__ inline cd csub(cd a, cd b) { return {a.r - b.r, a.i - b.i}; }
__host__ __device__ inline cd cmul(cd a, cd b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__host__ __device__ inline cd cconj(cd a) { return {a.r, -a.i}; }
__host__ __device__ inline double cabs2(cd a) { return a.r * a.r + a.i * a.i; }
__host__ __device__ inline cd cdiv(cd a, cd b) {
    const double d = cabs2(b);
    return {(a.r * b.r + a.i * b.i) / d, (a.i * b.r - a.r * b.i) / d};
}
__host__ __device__ inline void cross3(const cd *a, const cd *b, cd *o) {
    o[0] = csub(cmul(a[1], b[2]), cmul(a[2], b[1]));
    o[1] = csub(cmul(a[2], b[0]), cmul(a[0], b[2]));
    o[2] = csub(cmul(a[0], b[1]), cmul(a[1], b[0]));
}
__host__ __device__ inline double norm2_3(const cd *a) { return cabs2(a[0]) + cabs2(a[1]) + cabs2(a[2]); }

// eigenvector of the 3x3 Hermitian A for eigenvalue lam: conjugated cross product
// of two columns of A - lam I (ground_locg.py:506-513)
__host__ __device__ inline void vec3_at(const cd (*A)[3], double lam, cd *v) {
    cd col[3][3];
    double fro2 = 0.0;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            col[c][r] = (r == c) ? cd{A[r][c].r - lam, A[r][c].i} : A[r][c];
            fro2 += cabs2(col[c][r]);
        }
    cd w12[3], w01[3], w02[3];
    cross3(col[1], col[2], w12);
    cross3(col[0], col[1], w01);
    cross3(col[0], col[2], w02);
    const double n12 = norm2_3(w12), n01 = norm2_3(w01), n02 = norm2_3(w02);
    const double nmax = fmax(n12, fmax(n01, n02));
    cd w[3];
    double nw;
    if (nmax > 1e-28 * fro2 * fro2 && isfinite(nmax)) {
        // the reference's column pair (1, 2) unless it is ill-conditioned
        // relative to the best pair (then its direction is rounding noise)
        const cd *src = w12;
        nw = n12;
        if (!(n12 >= 1e-6 * nmax)) {
            if (n01 >= n02) {
                src = w01;
                nw = n01;
            } else {
                src = w02;
                nw = n02;
            }
        }
        for (int k = 0; k < 3; ++k) w[k] = src[k];
    } else {
        // eigenvalue of multiplicity >= 2: a vector orthogonal to the dominant column
        double best = -1.0;
        int bc = 0;
        for (int c = 0; c < 3; ++c) {
            const double nc = norm2_3(col[c]);
            if (nc > best) {
                best = nc;
                bc = c;
            }
        }
        if (best > 1e-28 * fro2 && best > 0.0) {
            int kmin = 0;
            for (int k = 1; k < 3; ++k)
                if (cabs2(col[bc][k]) < cabs2(col[bc][kmin])) kmin = k;
            cd e[3] = {{0, 0}, {0, 0}, {0, 0}};
            e[kmin] = {1.0, 0.0};
            cross3(col[bc], e, w);
        } else {
            w[0] = {1.0, 0.0};
            w[1] = {0.0, 0.0};
            w[2] = {0.0, 0.0};
        }
        nw = norm2_3(w);
    }
    const double inv = 1.0 / sqrt(nw);
    for (int k = 0; k < 3; ++k) v[k] = {w[k].r * inv, -w[k].i * inv};
}

__host__ __device__ inline double rayleigh(const cd (*A)[3], const cd *v, int n) {
    double acc = 0.0;
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) acc += cmul(cconj(v[r]), cmul(A[r][c], v[c])).r;
    return acc;
}

// eigenvector of the 2x2 Hermitian A for eigenvalue lam: the reference's mediant
// form [(d1 - lam + conj(b)) / (d0 - lam + b), -1] (ground_locg.py:474) unless its
// denominator is negligible, else the null vector of the larger row of A - lam I
__host__ __device__ inline void vec2_at(const cd (*A)[3], double lam, cd *v) {
    const double d0 = A[0][0].r, d1 = A[1][1].r;
    const cd b = A[1][0];
    const cd num = {d1 - lam + b.r, -b.i};
    const cd den = {d0 - lam + b.r, b.i};
    const double scale = (d0 - lam) * (d0 - lam) + (d1 - lam) * (d1 - lam) + cabs2(b);
    if (cabs2(den) > 1e-24 * scale && scale > 0.0) {
        const cd v0 = cdiv(num, den);
        const double nrm = sqrt(cabs2(v0) + 1.0);
        if (isfinite(nrm)) {
            v[0] = {v0.r / nrm, v0.i / nrm};
            v[1] = {-1.0 / nrm, 0.0};
            return;
        }
    }
    const cd a1 = {d1 - lam, 0.0}, mb = {-b.r, -b.i};
    const cd a2 = cconj(b), m0 = {-(d0 - lam), 0.0};
    cd w0 = a1, w1 = mb;
    if (cabs2(a2) + cabs2(m0) > cabs2(a1) + cabs2(mb)) {
        w0 = a2;
        w1 = m0;
    }
    const double nrm = sqrt(cabs2(w0) + cabs2(w1));
    if (nrm > 0.0) {
        v[0] = {w0.r / nrm, w0.i / nrm};
        v[1] = {w1.r / nrm, w1.i / nrm};
    } else {
        v[0] = {1.0, 0.0};
        v[1] = {0.0, 0.0};
    }
}

// Lowest eigenpair of a 2x2 / 3x3 Hermitian matrix m (row-major, complex
// interleaved); out = [lambda, 0, v0re, v0im, ...].
__host__ __device__ inline void eigen_small(const double *m, int n, double *out) {
    cd A[3][3];
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) A[r][c] = {m[2 * (r * n + c)], m[2 * (r * n + c) + 1]};
    double lam;
    cd v[3] = {{0, 0}, {0, 0}, {0, 0}};
    if (n == 2) {
        // ground_locg.py:467-478: eigenvalue from trace/determinant, mediant-form vector
        const double d0 = A[0][0].r, d1 = A[1][1].r;
        const double det = d0 * d1 - cabs2(A[1][0]);
        const double tr = d0 + d1;
        lam = (tr - sqrt(fmax(tr * tr - 4.0 * det, 0.0))) * 0.5;
        vec2_at(A, lam, v);
        for (int it = 0; it < 2; ++it) {   // Rayleigh-quotient refinement
            const double rq = rayleigh(A, v, 2);
            if (!isfinite(rq)) break;
            lam = rq;
            vec2_at(A, lam, v);
        }
    } else {
        // ground_locg.py:481-514 (J. Kopp, Int. J. Mod. Phys. C 19, 523 (2008)):
        // characteristic polynomial, trigonometric root, eigenvector as the
        // conjugated cross product of two columns of A - lam I
        const double d[3] = {A[0][0].r, A[1][1].r, A[2][2].r};
        const double mo[3] = {cabs2(A[1][0]), cabs2(A[2][0]), cabs2(A[2][1])};
        const double c2 = -(d[0] + d[1] + d[2]);
        const double c1 = (d[0] * d[2] + d[1] * d[0] + d[2] * d[1]) - (mo[0] + mo[1] + mo[2]);
        double c0 = d[0] * mo[2] + d[1] * mo[1] + d[2] * mo[0];
        c0 -= d[0] * d[1] * d[2];
        c0 -= 2.0 * cmul(cmul(A[0][2], A[1][0]), A[2][1]).r;
        const double p = c2 * c2 - 3.0 * c1;
        const double q = -13.5 * c0 - c2 * c2 * c2 + 4.5 * c2 * c1;
        const double phi =
            atan2(sqrt(fmax(0.0, 27.0 * (0.25 * c1 * c1 * (p - c1) + c0 * (q + 6.75 * c0)))), q) / 3.0;
        const double cphi = cos(phi), sphi = sin(phi), s3 = sqrt(3.0);
        const double xmin = fmin(fmin(2.0 * cphi, -cphi - s3 * sphi), -cphi + s3 * sphi);
        lam = sqrt(fmax(p, 0.0)) / 3.0 * xmin - c2 / 3.0;
        vec3_at(A, lam, v);
        for (int it = 0; it < 2; ++it) {   // Rayleigh-quotient refinement
            const double rq = rayleigh(A, v, 3);
            if (!isfinite(rq)) break;
            lam = rq;
            vec3_at(A, lam, v);
        }
    }
    out[0] = lam;
    out[1] = 0.0;
    for (int k = 0; k < n; ++k) {
        out[2 + 2 * k] = v[k].r;
        out[3 + 2 * k] = v[k].i;
    }
}

__global__ void k_eigen_small(const double *__restrict__ m, int n, double *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) eigen_small(m, n, out);
}

// ---------------------------------------------------------------- batched vector kernels
// Vectors of a batched plan are the concatenation of every problem's rows
// (rows[k] .. rows[k+1]); these kernels run one LOCG vector step for all
// problems at once.  Grid: (chunks, problems).
constexpr int MAXV = 9;

struct PtrPack {
    const void *a[MAXV];
    const void *b[MAXV];
};

__device__ inline void load_c(const double &v, double &re, double &im) { re = v; im = 0.0; }
__device__ inline void load_c(const float &v, double &re, double &im) { re = v; im = 0.0; }
__device__ inline void load_c(const double2 &v, double &re, double &im) { re = v.x; im = v.y; }
__device__ inline void load_c(const float2 &v, double &re, double &im) { re = v.x; im = v.y; }
__device__ inline void store_c(double &o, double re, double) { o = re; }
__device__ inline void store_c(float &o, double re, double) { o = static_cast<float>(re); }
__device__ inline void store_c(double2 &o, double re, double im) { o = make_double2(re, im); }
__device__ inline void store_c(float2 &o, double re, double im) { o = make_float2(static_cast<float>(re), static_cast<float>(im)); }

// Per-block partial sums of conj(a_q[i]) * b_q[i] over the rows i of problem k:
// part[(k * gridDim.x + blockIdx.x) * 2 * NP + {2q, 2q + 1}].  Every block writes
// its slot (zeros when it has no rows); k_dots_finish adds the slots in block
// order, so the sums do not depend on the order in which blocks run.
template <typename V, int NP>
__global__ __launch_bounds__(256) void k_batch_dots(PtrPack pk, const int64_t *__restrict__ rows,
                                                    double *__restrict__ part) {
    __shared__ double s_red[4][2 * NP];
    const int k = blockIdx.y;
    const int64_t r0 = rows[k], r1 = rows[k + 1];
    double re[NP], im[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) re[q] = im[q] = 0.0;
    for (int64_t i = r0 + static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < r1;
         i += static_cast<int64_t>(gridDim.x) * 256) {
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            double ar, ai, br, bi;
            load_c(static_cast<const V *>(pk.a[q])[i], ar, ai);
            load_c(static_cast<const V *>(pk.b[q])[i], br, bi);
            re[q] += ar * br + ai * bi;
            im[q] += ar * bi - ai * br;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        const double sr = wave_sum(re[q]), si = wave_sum(im[q]);
        if (lane == 0) {
            s_red[wave][2 * q] = sr;
            s_red[wave][2 * q + 1] = si;
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * NP) {
        const double v = s_red[0][threadIdx.x] + s_red[1][threadIdx.x] + s_red[2][threadIdx.x] + s_red[3][threadIdx.x];
        part[(static_cast<int64_t>(k) * gridDim.x + blockIdx.x) * (2 * NP) + threadIdx.x] = v;
    }
}

// Sum of the `chunks` partial slots of problem k (slot stride `width` doubles, `n`
// values each) in a fixed order: lane l adds slots l, l + 64, ... in turn, then the
// butterfly sum over the wave.  Lane 0 receives value t in its return.
__device__ inline double part_sum(const double *__restrict__ part, int k, int chunks, int width, int t) {
    const int lane = threadIdx.x & 63;
    double acc = 0.0;
    for (int c = lane; c < chunks; c += 64) acc += part[(static_cast<int64_t>(k) * chunks + c) * width + t];
    return wave_sum(acc);
}

// out[2 (k NP + q) + {0, 1}] = the sums of problem k's partial slots; one wave per problem.
__global__ __launch_bounds__(64) void k_dots_finish(const double *__restrict__ part, int chunks, int np,
                                                   double *__restrict__ out) {
    const int k = blockIdx.x;
    for (int t = 0; t < 2 * np; ++t) {
        const double v = part_sum(part, k, chunks, 2 * np, t);
        if (threadIdx.x == 0) out[2 * static_cast<int64_t>(k) * np + t] = v;
    }
}

// out[i] = sum_q coef[k * NV + q] * v_q[i] for the rows i of problem k (coef complex double;
// only the real part is used for real vectors).  out may be one of the inputs.
template <typename V, int NV>
__global__ __launch_bounds__(256) void k_batch_combine(PtrPack pk, const int64_t *__restrict__ rows,
                                                       const double *__restrict__ coef, void *out) {
    const int k = blockIdx.y;
    const int64_t r0 = rows[k], r1 = rows[k + 1];
    double cr[NV], ci[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        cr[q] = coef[2 * (static_cast<int64_t>(k) * NV + q)];
        ci[q] = coef[2 * (static_cast<int64_t>(k) * NV + q) + 1];
    }
    for (int64_t i = r0 + static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < r1;
         i += static_cast<int64_t>(gridDim.x) * 256) {
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double vr, vi;
            load_c(static_cast<const V *>(pk.a[q])[i], vr, vi);
            sr += cr[q] * vr - ci[q] * vi;
            si += cr[q] * vi + ci[q] * vr;
        }
        store_c(static_cast<V *>(out)[i], sr, si);
    }
}

// ---------------------------------------------------------------- fused LOCG step
// One pass over the rows of every problem: outputs o = sum_q coef[o][q] in_q
// (per-problem complex coefficients, written by the previous scalar stage),
// and complex dots conj(u) v over inputs and outputs accumulated per problem.
// Outputs may alias inputs (each element is read before it is written).  With
// LO_IN the last input is stored in the low precision type LV; with LO_OUT the
// last output is.
constexpr int LP_IN = 6, LP_OUT = 4, LP_DOT = 9;

struct LPass {
    const void *in[LP_IN];
    void *out[LP_OUT];
    const double *coef;      // problem k: coef + k * stride, [o][q] complex
    double *part;            // per-block partial dots: (k * gridDim.x + block) * 2 LP_DOT doubles
    int32_t stride;          // doubles of per-problem work space
};

// The dot products of a pass, as a compile-time list: pair d is operands
// P_d / 16 and P_d % 16 (0..NIN-1 the inputs, NIN + o output o), conj(first) * second.
template <int... P> struct Dots {
    static constexpr int n = sizeof...(P);
    static constexpr int pairs[n > 0 ? n : 1] = {P...};
};

// re, im as the vector type V stores them: FP32 rounding, no imaginary part for
// real vectors (the dots of a pass use the values its outputs hold)
__device__ inline void as_stored(double &, double &im, const double *) { im = 0.0; }
__device__ inline void as_stored(double &re, double &im, const float *) {
    re = static_cast<double>(static_cast<float>(re));
    im = 0.0;
}
__device__ inline void as_stored(double &, double &, const double2 *) {}
__device__ inline void as_stored(double &re, double &im, const float2 *) {
    re = static_cast<double>(static_cast<float>(re));
    im = static_cast<double>(static_cast<float>(im));
}

template <typename V, typename LV, int NIN, int NOUT, typename DL, bool LO_IN, bool LO_OUT>
__global__ __launch_bounds__(256) void k_locg_pass(LPass a, const int64_t *__restrict__ rows) {
    constexpr int NDOT = DL::n;
    static_assert(NDOT <= LP_DOT, "dots per pass");
    const int k = blockIdx.y;
    const int64_t r0 = rows[k], r1 = rows[k + 1];
    const double *cf = a.coef + static_cast<int64_t>(k) * a.stride;
    double cr[NOUT > 0 ? NOUT : 1][NIN], ci[NOUT > 0 ? NOUT : 1][NIN];
#pragma unroll
    for (int o = 0; o < NOUT; ++o)
#pragma unroll
        for (int q = 0; q < NIN; ++q) {
            cr[o][q] = cf[2 * (o * NIN + q)];
            ci[o][q] = cf[2 * (o * NIN + q) + 1];
        }
    double dre[NDOT > 0 ? NDOT : 1], dim_[NDOT > 0 ? NDOT : 1];
#pragma unroll
    for (int d = 0; d < NDOT; ++d) dre[d] = dim_[d] = 0.0;
    for (int64_t i = r0 + static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < r1;
         i += static_cast<int64_t>(gridDim.x) * 256) {
        double vr[NIN + NOUT], vi[NIN + NOUT];
#pragma unroll
        for (int q = 0; q < NIN; ++q) {
            if (LO_IN && q == NIN - 1) load_c(static_cast<const LV *>(a.in[q])[i], vr[q], vi[q]);
            else load_c(static_cast<const V *>(a.in[q])[i], vr[q], vi[q]);
        }
#pragma unroll
        for (int o = 0; o < NOUT; ++o) {
            double sr = 0.0, si = 0.0;
#pragma unroll
            for (int q = 0; q < NIN; ++q) {
                sr += cr[o][q] * vr[q] - ci[o][q] * vi[q];
                si += cr[o][q] * vi[q] + ci[o][q] * vr[q];
            }
            if (LO_OUT && o == NOUT - 1) {
                store_c(static_cast<LV *>(a.out[o])[i], sr, si);
            } else {
                as_stored(sr, si, static_cast<const V *>(nullptr));
                store_c(static_cast<V *>(a.out[o])[i], sr, si);
            }
            vr[NIN + o] = sr;
            vi[NIN + o] = si;
        }
#pragma unroll
        for (int d = 0; d < NDOT; ++d) {
            const int pa = DL::pairs[d] / 16, pb = DL::pairs[d] % 16;
            const double ar = vr[pa], ai = vi[pa], br = vr[pb], bi = vi[pb];
            dre[d] += ar * br + ai * bi;
            dim_[d] += ar * bi - ai * br;
        }
    }
    if constexpr (NDOT > 0) {
        __shared__ double s_red[4][2 * LP_DOT];
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
        for (int d = 0; d < NDOT; ++d) {
            const double sr = wave_sum(dre[d]), si = wave_sum(dim_[d]);
            if (lane == 0) {
                s_red[wave][2 * d] = sr;
                s_red[wave][2 * d + 1] = si;
            }
        }
        __syncthreads();
        if (threadIdx.x < 2 * NDOT) {
            const double v = s_red[0][threadIdx.x] + s_red[1][threadIdx.x] + s_red[2][threadIdx.x] +
                             s_red[3][threadIdx.x];
            a.part[(static_cast<int64_t>(k) * gridDim.x + blockIdx.x) * (2 * LP_DOT) + threadIdx.x] = v;
        }
    }
}

// Per-problem work space of the fused step (doubles, complex slots at 2 * slot):
// dot slots 0..19, coefficient blocks after them.
enum {
    LW_D_XAX = 0, LW_D_RR = 1, LW_D_XR = 2, LW_D_YR = 3, LW_D_XP = 4, LW_D_YP = 5, LW_D_PP = 6,
    LW_D_RR9 = 7, LW_D_XX = 16, LW_D_XY = 17, LW_D_YY = 18, LW_D_YY2 = 19,
    LW_C2 = 20, LW_C3 = 23, LW_C4 = 26, LW_C5 = 32, LW_C6 = 38, LW_C7 = 62, LW_C8 = 78,
    LW_SLOTS = 82
};
constexpr int LOCG_WORK = 2 * LW_SLOTS + 4;   // doubles per problem
static_assert(LOCG_WORK == RQX_LOCG_WORK, "work size in rqx.h");
static_assert(RQX_LOCG_SCALE >= 2 * LW_SLOTS && RQX_LOCG_SCALE < LOCG_WORK, "scale slot in rqx.h");

__device__ inline void put_c(double *w, int slot, double re, double im) {
    w[2 * slot] = re;
    w[2 * slot + 1] = im;
}

// Scalar stages between the passes: one wave per problem.  The wave first adds the
// preceding pass's `ndot` partial dots (fixed order, part_sum) into the work slots
// from `dot_slot` on; thread 0 then runs the stage.
__global__ __launch_bounds__(64) void k_locg_scalar(int stage, int K, double *__restrict__ work,
                                                    double *__restrict__ theta, double *__restrict__ rel,
                                                    uint8_t *__restrict__ active, double tol, int real,
                                                    const double *__restrict__ part, int chunks, int dot_slot,
                                                    int ndot) {
    const int k = blockIdx.x;
    if (k >= K) return;
    double *w = work + static_cast<int64_t>(k) * LOCG_WORK;
    for (int t = 0; t < 2 * ndot; ++t) {
        const double v = part_sum(part, k, chunks, 2 * LP_DOT, t);
        if (threadIdx.x == 0) w[2 * dot_slot + t] = v;
    }
    if (threadIdx.x != 0) return;
    auto re = [&](int slot) { return w[2 * slot]; };
    auto im = [&](int slot) { return w[2 * slot + 1]; };
    if (stage == 1) {                  // theta = <x, Ax>;  r = Ax - theta x
        const double th = re(LW_D_XAX);
        theta[k] = th;
        put_c(w, LW_C2 + 0, 1.0, 0.0);
        put_c(w, LW_C2 + 1, -th, 0.0);
        put_c(w, LW_C2 + 2, 0.0, 0.0);
    } else if (stage == 2) {           // residual norm, freeze; first Gram-Schmidt pass
        // relative to max(|theta|, the caller's scale): still defined where the eigenvalue is 0
        const double rn = sqrt(fmax(re(LW_D_RR), 0.0));
        const double rl = rn / fmax(fmax(fabs(theta[k]), w[RQX_LOCG_SCALE]), 1e-300);
        rel[k] = rl;
        if (!(rl > tol)) active[k] = 0;
        put_c(w, LW_C3 + 0, 1.0, 0.0);
        put_c(w, LW_C3 + 1, -re(LW_D_XR), -im(LW_D_XR));
        put_c(w, LW_C3 + 2, -re(LW_D_YR), -im(LW_D_YR));
    } else if (stage == 3) {           // second pass, normalised (both the vector and its low copy)
        const double ar = re(LW_D_XP), ai = im(LW_D_XP), br = re(LW_D_YP), bi = im(LW_D_YP);
        const double n2 = re(LW_D_PP) - (ar * ar + ai * ai) - (br * br + bi * bi);
        const double inv = n2 > 0.0 ? 1.0 / sqrt(n2) : 0.0;
        for (int o = 0; o < 2; ++o) {
            put_c(w, LW_C4 + 3 * o + 0, inv, 0.0);
            put_c(w, LW_C4 + 3 * o + 1, -ar * inv, -ai * inv);
            put_c(w, LW_C4 + 3 * o + 2, -br * inv, -bi * inv);
        }
        for (int q = 0; q < 6; ++q) put_c(w, LW_C5 + q, q == 5 ? 1.0 : 0.0, 0.0);
    } else if (stage == 5) {           // Rayleigh-Ritz on [x, y, p] -> coefficients
        double m[18], out[8];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                const int a = LW_D_RR9 + 3 * r + c, b = LW_D_RR9 + 3 * c + r;
                m[2 * (3 * r + c)] = 0.5 * (re(a) + re(b));
                m[2 * (3 * r + c) + 1] = 0.5 * (im(a) - im(b));
            }
        eigen_small(m, 3, out);
        double c0r = out[2], c0i = out[3], c1r = out[4], c1i = out[5], c2r = out[6], c2i = out[7];
        if (real) c0i = c1i = c2i = 0.0;
        if (!active[k]) {
            c0r = 1.0;
            c0i = c1r = c1i = c2r = c2i = 0.0;
        }
        for (int q = 0; q < 24; ++q) put_c(w, LW_C6 + q, 0.0, 0.0);
        // inputs (x, ax, y, ay, p, ap) -> outputs (y, ay, x, ax)
        put_c(w, LW_C6 + 0 * 6 + 2, c1r, c1i);
        put_c(w, LW_C6 + 0 * 6 + 4, c2r, c2i);
        put_c(w, LW_C6 + 1 * 6 + 3, c1r, c1i);
        put_c(w, LW_C6 + 1 * 6 + 5, c2r, c2i);
        put_c(w, LW_C6 + 2 * 6 + 0, c0r, c0i);
        put_c(w, LW_C6 + 2 * 6 + 2, c1r, c1i);
        put_c(w, LW_C6 + 2 * 6 + 4, c2r, c2i);
        put_c(w, LW_C6 + 3 * 6 + 1, c0r, c0i);
        put_c(w, LW_C6 + 3 * 6 + 3, c1r, c1i);
        put_c(w, LW_C6 + 3 * 6 + 5, c2r, c2i);
    } else if (stage == 6) {           // normalise x; y minus its projection on x
        const double nx2 = re(LW_D_XX);
        const double inv = nx2 > 0.0 ? 1.0 / sqrt(nx2) : 0.0;
        const double pr = nx2 > 0.0 ? re(LW_D_XY) / nx2 : 0.0, pi = nx2 > 0.0 ? im(LW_D_XY) / nx2 : 0.0;
        for (int q = 0; q < 16; ++q) put_c(w, LW_C7 + q, 0.0, 0.0);
        // inputs (x, ax, y, ay) -> outputs (x, ax, y, ay)
        put_c(w, LW_C7 + 0 * 4 + 0, inv, 0.0);
        put_c(w, LW_C7 + 1 * 4 + 1, inv, 0.0);
        put_c(w, LW_C7 + 2 * 4 + 0, -pr, -pi);
        put_c(w, LW_C7 + 2 * 4 + 2, 1.0, 0.0);
        put_c(w, LW_C7 + 3 * 4 + 1, -pr, -pi);
        put_c(w, LW_C7 + 3 * 4 + 3, 1.0, 0.0);
    } else if (stage == 7) {           // normalise y and its image
        const double ny2 = re(LW_D_YY2);
        const double inv = ny2 > 0.0 ? 1.0 / sqrt(ny2) : 0.0;
        put_c(w, LW_C8 + 0, inv, 0.0);
        put_c(w, LW_C8 + 1, 0.0, 0.0);
        put_c(w, LW_C8 + 2, 0.0, 0.0);
        put_c(w, LW_C8 + 3, inv, 0.0);
    }
}

// entries per batch of the cached matvec: the column and value loads of a batch,
// then its gathers of x, are all in flight together (16 while a batch's
// registers stay small, 8 for 16-byte vector entries or complex values)
template <typename V, typename CV> constexpr int spmv_unroll() {
#ifdef SPMV_UNROLL
    return SPMV_UNROLL;
#else
    return (sizeof(V) + sizeof(CV) <= 12) ? 16 : 8;
#endif
}

// workgroup -> 256-row block of the cached matvec: contiguous blocks per XCD, so
// one XCD's L2 sees the gathers of neighbouring rows
__device__ inline uint32_t spmv_block() { return xcd_remap(blockIdx.x, gridDim.x); }

// The cached matvec's workgroups take SPMV_WG rows (8 slices) each and first copy
// x over their staging window -- the rows whose states share the high part (bits
// >= the problem's section split) of the workgroup's middle row, clipped to the
// LDS budget around that row -- into LDS.  Second-section hits whose source lies
// in the window (in a product subspace: the flips of low strings only, the bulk
// of them) then read x from LDS; the others read global memory.  The values read
// are the same either way.
constexpr int SPMV_WG = 512;
constexpr int SPMV_LDS_BYTES = 64 * 1024;

template <typename V> constexpr int stage_cap() { return SPMV_LDS_BYTES / static_cast<int>(sizeof(V)); }

// x[c] from the staged window [lo, lo + len) when it holds c, else from memory;
// branch free (both loads issue; a lane outside the window reads the window's
// first entry from memory, the same address for every such lane)
template <typename V>
__device__ inline V gather_staged(const V *__restrict__ x, const V *xs, int32_t lo, int32_t len, int32_t c) {
    const uint32_t o = static_cast<uint32_t>(c - lo);
    const bool in = o < static_cast<uint32_t>(len);
    V a = xs[in ? o : 0u];
    if (!in) a = x[c];   // a global load for the lanes outside the window only
    return a;
}

template <typename T> __device__ inline T stream_load(const T *p) { return __builtin_nontemporal_load(p); }
__device__ inline float2 stream_load(const float2 *p) {
    return make_float2(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y));
}
__device__ inline double2 stream_load(const double2 *p) {
    return make_double2(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y));
}

// one packet (PACK consecutive entries of a row) with 16-byte non-temporal
// loads; p: the row's first entry of the packet (its 16-byte pieces are 64 e
// entries apart)
template <typename T>
__device__ inline void load_packet(const T *p, T *out) {
    constexpr int E = pack_sub<T>();
    static_assert(sizeof(T) * E == 16, "16-byte pieces");
    constexpr int N = PACK / E;
    i32x4 t[N];
#pragma unroll
    for (int i = 0; i < N; ++i) t[i] = __builtin_nontemporal_load(reinterpret_cast<const i32x4 *>(p + i * SLICE * E));
    __builtin_memcpy(out, t, sizeof(t));
}

// acc += the w entries (a multiple of PACK) of one section of a row, in order:
// column and value packets of U entries, then their gathers of x, then the
// FMAs.  c, vv: the row's first packet.
template <int U, typename V, typename CV>
__device__ inline void spmv_section(V &acc, const int32_t *c, const CV *vv, uint32_t w, const V *__restrict__ x) {
    static_assert(U % PACK == 0, "whole packets per batch");
    constexpr int64_t STEP = SLICE * PACK;        // entries between a row's packets
    uint32_t k = 0;
    for (; k + U <= w; k += U) {
        int32_t ci[U];
        CV cv[U];
        V xv[U];
#pragma unroll
        for (int pk = 0; pk < U / PACK; ++pk) {
            load_packet(c + pk * STEP, ci + pk * PACK);
            load_packet(vv + pk * STEP, cv + pk * PACK);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x[ci[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) fma_acc(acc, cv[u], xv[u]);
        c += (U / PACK) * STEP;
        vv += (U / PACK) * STEP;
    }
    for (; k < w; k += PACK) {
        int32_t ci[PACK];
        CV cv[PACK];
        load_packet(c, ci);
        load_packet(vv, cv);
#pragma unroll
        for (int u = 0; u < PACK; ++u) fma_acc(acc, cv[u], x[ci[u]]);
        c += STEP;
        vv += STEP;
    }
}

// one packet of a SELL level: its 16-byte pieces are n e entries apart
template <typename T>
__device__ inline void load_packet_sell(const T *p, int32_t n, T *out) {
    constexpr int E = pack_sub<T>();
    static_assert(sizeof(T) * E == 16, "16-byte pieces");
    constexpr int N = PACK / E;
    i32x4 t[N];
#pragma unroll
    for (int i = 0; i < N; ++i)
        t[i] = __builtin_nontemporal_load(reinterpret_cast<const i32x4 *>(p + static_cast<int64_t>(i) * n * E));
    __builtin_memcpy(out, t, sizeof(t));
}

// acc += section B of the row at SELL position pos, in order: level q holds a
// packet for the positions below n_q (the chunk widths W are wave-uniform, so the
// level bases are scalar), U / PACK levels' loads, then their gathers, then the
// multiply-adds.  Branch free: a position past n_q reads position 0's packet
// (the same cache lines) and adds value 0 times its own row's x, as the padding
// of a fully padded layout would.
template <int U, typename V, typename CV>
__device__ inline void spmv_section_sell(V &acc, const int32_t *c, const CV *vv, const uint32_t *W, int pos,
                                         int32_t row, const V *__restrict__ x, const V *xs, int32_t lo, int32_t len) {
    static_assert(U % PACK == 0, "whole packets per batch");
    constexpr int L = U / PACK;
    constexpr int EC = pack_sub<int32_t>(), EV = pack_sub<CV>();
    const CV zero = vzero(static_cast<CV *>(nullptr));
    const uint32_t w0 = W[0];
    uint32_t q = 0;
    for (; q + L <= w0; q += L) {
        int32_t ci[U];
        CV cv[U];
        V xv[U];
        bool on[L];
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const SellLevel lv = sell_level(W, q + l);
            on[l] = pos < lv.n;
            const int at = on[l] ? pos : 0;
            load_packet_sell(c + lv.base + at * EC, lv.n, ci + l * PACK);
            load_packet_sell(vv + lv.base + at * EV, lv.n, cv + l * PACK);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ci[u] = on[u / PACK] ? ci[u] : row;
            cv[u] = on[u / PACK] ? cv[u] : zero;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = gather_staged(x, xs, lo, len, ci[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) fma_acc(acc, cv[u], xv[u]);
    }
    for (; q < w0; ++q) {
        const SellLevel lv = sell_level(W, q);
        const bool on = pos < lv.n;
        const int at = on ? pos : 0;
        int32_t ci[PACK];
        CV cv[PACK];
        load_packet_sell(c + lv.base + at * EC, lv.n, ci);
        load_packet_sell(vv + lv.base + at * EV, lv.n, cv);
#pragma unroll
        for (int u = 0; u < PACK; ++u)
            fma_acc(acc, on ? cv[u] : zero, gather_staged(x, xs, lo, len, on ? ci[u] : row));
    }
}

// a wave-uniform 64-bit value, held in scalar registers
__device__ inline int64_t uniform64(int64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(static_cast<uint64_t>(v) >> 32));
    return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

// section-A level whose columns are not row + (one offset for the whole slice)
constexpr int32_t LEVEL_GATHER = INT32_MIN;

// acc += the w section-A entries of a row, in order.  Level k of the slice is
// affine (del[k] = d: every lane's source row is its own row + d, so the
// column is computed and the gather is contiguous) or stored (del[k] =
// LEVEL_GATHER: columns read from c, 64 lane-contiguous per such level); and
// has one value for all lanes (code[k] = -1: aval[k], a scalar load) or 64
// (code[k] = j, the count of such levels before k: read from va + 64 j,
// lane-contiguous).  Returns the number of stored-column levels.
template <int U, typename V, typename CV>
__device__ inline uint32_t spmv_section_a(V &acc, const int32_t *c, const CV *va, const CV *__restrict__ aval,
                                          const int32_t *__restrict__ del, const int32_t *__restrict__ code,
                                          uint32_t w, int32_t row, const V *__restrict__ x) {
    uint32_t k = 0, j = 0;
    for (; k + U <= w; k += U) {
        int32_t d[U], cd[U], ci[U];
        CV cv[U];
        V xv[U];
        // codes and one-value levels' values are wave-uniform: scalar loads (every
        // level has an aval slot); only stored-value levels issue a vector load
#pragma unroll
        for (int u = 0; u < U; ++u) {
            d[u] = ldc(del + k + u);
            cd[u] = ldc(code + k + u);
            cv[u] = ldc(aval + k + u);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (cd[u] >= 0) cv[u] = stream_load(va + static_cast<int64_t>(cd[u]) * SLICE);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (d[u] == LEVEL_GATHER) {
                ci[u] = stream_load(c + j * SLICE);
                ++j;
            } else {
                ci[u] = row + d[u];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x[ci[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) fma_acc(acc, cv[u], xv[u]);
    }
    for (; k < w; ++k) {
        const int32_t d = ldc(del + k), cd = ldc(code + k);
        CV cv = ldc(aval + k);
        if (cd >= 0) cv = stream_load(va + static_cast<int64_t>(cd) * SLICE);
        int32_t ci = row + d;
        if (d == LEVEL_GATHER) {
            ci = stream_load(c + j * SLICE);
            ++j;
        }
        fma_acc(acc, cv, x[ci]);
    }
    return j;
}

// y = H x from the hit cache: one wave per 64-row slice, one row per lane,
// entries read lane-contiguously (and marked non-temporal: they are read once
// per matvec, x is the data worth keeping in L2); the same accumulation
// sequence as k_tiles (diagonal and section B, section A, their sum), so
// results are identical.  Per slice s: section-A level codes (offsets, value
// codes) and single values at aoff[s], the other section-A values at avoff[s],
// stored columns at coff[s] (stored section-A levels, then B), section-B
// values at boff[s].
template <typename R, bool CPLXV, bool CPLXC>
__global__ __launch_bounds__(SPMV_WG) void k_spmv_sell(const int32_t *__restrict__ col, const void *__restrict__ valb,
                                                       const void *__restrict__ vala, const void *__restrict__ aval,
                                                       const int64_t *__restrict__ boff, const int64_t *__restrict__ avoff,
                                                       const int64_t *__restrict__ coff, const int64_t *__restrict__ aoff,
                                                       const int32_t *__restrict__ adel, const int32_t *__restrict__ acode,
                                                       const uint32_t *__restrict__ wa, const uint8_t *__restrict__ bperm,
                                                       const uint16_t *__restrict__ bw, const int2 *__restrict__ win,
                                                       const void *__restrict__ diag, int64_t n_states,
                                                       const void *__restrict__ xin, void *__restrict__ yout) {
    using V = typename VT<R, CPLXV>::T;
    using CC = typename C2<R>::T;
    using CV = typename std::conditional<CPLXC, CC, R>::type;
    constexpr int U = spmv_unroll<V, CV>();
    extern __shared__ int4 stage_lds[];
    V *xs = reinterpret_cast<V *>(stage_lds);
    const V *x = static_cast<const V *>(xin);
    const uint32_t wg = spmv_block();
    // the staging window: the workgroup's block of rows, clipped to the LDS budget
    // around its middle row
    const int2 bwin = ldc(win + wg);
    int32_t lo = bwin.x, len = bwin.y;
    if (len > stage_cap<V>()) {
        const int64_t mid = std::min<int64_t>(static_cast<int64_t>(wg) * SPMV_WG + SPMV_WG / 2, n_states - 1);
        const int64_t l = std::max<int64_t>(lo, std::min<int64_t>(mid - stage_cap<V>() / 2,
                                                                  static_cast<int64_t>(lo) + len - stage_cap<V>()));
        lo = static_cast<int32_t>(l);
        len = stage_cap<V>();
    }
    for (int32_t i = threadIdx.x; i < len; i += SPMV_WG) xs[i] = x[lo + i];
    __syncthreads();
    // a wave's 64 lanes are one slice; lane l takes the slice's row at SELL position l
    const int64_t first = static_cast<int64_t>(wg) * SPMV_WG + (threadIdx.x & ~(SLICE - 1));
    const int64_t slice = uniform64(first / SLICE);   // per-slice tables: scalar loads
    const int pos = static_cast<int>(threadIdx.x % SLICE);
    if (first >= n_states) return;
    const int lane = bperm[slice * SLICE + pos];      // the row within the slice
    const int64_t row = slice * SLICE + lane;
    if (row >= n_states) return;
    V acc = vzero(static_cast<V *>(nullptr));
    if (diag) fma_acc(acc, static_cast<const CC *>(diag)[row], x[row]);   // NULL: no problem has an x = 0 group
    const uint32_t w_a = ldc(wa + slice);
    const int64_t bo = ldc(boff + slice), avo = ldc(avoff + slice);
    const int64_t co = ldc(coff + slice), ao = ldc(aoff + slice);
    uint32_t W[NCH];
    const i32x4 wv = ldc(reinterpret_cast<const i32x4 *>(bw + slice * NCH));
#pragma unroll
    for (int j = 0; j < NCH / 2; ++j) {
        W[2 * j] = static_cast<uint32_t>(wv[j]) & 0xffffu;
        W[2 * j + 1] = static_cast<uint32_t>(wv[j]) >> 16;
    }
    V acc_a = vzero(static_cast<V *>(nullptr));
    const uint32_t n_stored = spmv_section_a<U>(acc_a, col + co + lane, static_cast<const CV *>(vala) + avo + lane,
                                                static_cast<const CV *>(aval) + ao, adel + ao, acode + ao, w_a,
                                                static_cast<int32_t>(row), x);
    spmv_section_sell<U>(acc, col + co + static_cast<int64_t>(n_stored) * SLICE,
                         static_cast<const CV *>(valb) + bo, W, pos, static_cast<int32_t>(row), x, xs, lo, len);
    static_cast<V *>(yout)[row] = vadd(acc, acc_a);
}

// Staging windows of the cached matvec (one thread per workgroup of SPMV_WG rows):
// the rows around the workgroup's middle row whose states agree with it on every
// bit at or above its problem's section split (a_bits >= 64: the whole problem).
// win[w] = (first row, rows); the kernel clips it to its LDS budget.
__global__ void k_stage_windows(const uint64_t *__restrict__ keys, const int64_t *__restrict__ prob_rows,
                                const ProbDesc *__restrict__ prob, int K, int64_t n_states, int64_t n_wg,
                                int2 *__restrict__ win, int *__restrict__ win_max) {
    const int64_t w = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (w >= n_wg) return;
    const int64_t r = std::min<int64_t>(w * SPMV_WG + SPMV_WG / 2, n_states - 1);
    int lo_k = 0, hi_k = K;                  // the problem of row r: prob_rows[k] <= r < prob_rows[k + 1]
    while (hi_k - lo_k > 1) {
        const int m = (lo_k + hi_k) / 2;
        if (prob_rows[m] <= r) lo_k = m;
        else hi_k = m;
    }
    const int64_t p0 = prob_rows[lo_k], p1 = prob_rows[lo_k + 1];
    const int a = prob[lo_k].a_bits;
    const uint64_t hi = a >= 64 ? 0ull : (keys[r] >> a);
    auto high = [&](int64_t i) { return a >= 64 ? 0ull : (keys[i] >> a); };
    int64_t b = p0, e = r;                   // first row of the block in [p0, r]
    while (b < e) {
        const int64_t m = b + (e - b) / 2;
        if (high(m) < hi) b = m + 1;
        else e = m;
    }
    int64_t b2 = r + 1, e2 = p1;             // one past its last row in [r + 1, p1]
    while (b2 < e2) {
        const int64_t m = b2 + (e2 - b2) / 2;
        if (high(m) <= hi) b2 = m + 1;
        else e2 = m;
    }
    const int64_t len = std::min<int64_t>(b2 - b, INT32_MAX);
    win[w] = make_int2(static_cast<int32_t>(b), static_cast<int32_t>(len));
    atomicMax(win_max, static_cast<int>(len));
}

// Section widths of each slice (one wave per slice): the largest row count of
// each section over the slice's rows, rounded up to whole packets; the hits of
// all rows are summed into nnz[0] and the section-B hits into nnz[1] (one atomic
// each per wave).  Section B's SELL order: each row's position (descending
// section-B count, ties in row order; rows past the states count 0) in bpos,
// the row of each position in bperm, the chunk widths (packets) in bw and their
// sum in bsum.
__global__ __launch_bounds__(256) void k_slice_widths(const uint32_t *__restrict__ cnt_b,
                                                      const uint32_t *__restrict__ cnt_a, int64_t n_states,
                                                      int64_t n_slices, uint32_t *__restrict__ wa,
                                                      uint32_t *__restrict__ wb, uint32_t *__restrict__ bsum,
                                                      uint8_t *__restrict__ bpos, uint8_t *__restrict__ bperm,
                                                      uint16_t *__restrict__ bw, unsigned long long *nnz) {
    const int64_t slice = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / SLICE;
    if (slice >= n_slices) return;
    const int lane = static_cast<int>(threadIdx.x % SLICE);
    const int64_t row = slice * SLICE + lane;
    const uint32_t my_b = row < n_states ? cnt_b[row] : 0u;
    uint32_t b = my_b, a = row < n_states ? cnt_a[row] : 0u;
    unsigned long long sum = static_cast<unsigned long long>(a) + b, sum_b = b;
    for (int off = 32; off > 0; off >>= 1) {
        b = max(b, static_cast<uint32_t>(__shfl_xor(static_cast<int>(b), off, 64)));
        a = max(a, static_cast<uint32_t>(__shfl_xor(static_cast<int>(a), off, 64)));
        sum += __shfl_xor(sum, off, 64);
        sum_b += __shfl_xor(sum_b, off, 64);
    }
    int rank = 0;
    for (int m = 0; m < SLICE; ++m) {
        const uint32_t c = static_cast<uint32_t>(__shfl(static_cast<int>(my_b), m, 64));
        rank += (c > my_b || (c == my_b && m < lane)) ? 1 : 0;
    }
    bpos[row] = static_cast<uint8_t>(rank);
    bperm[slice * SLICE + rank] = static_cast<uint8_t>(lane);
    const uint32_t w_chunk = (my_b + PACK - 1) / PACK;       // the widest row of a chunk is its first position
    if (rank % CH == 0) bw[slice * NCH + rank / CH] = static_cast<uint16_t>(w_chunk);
    uint32_t ws = rank % CH == 0 ? w_chunk : 0u;
    for (int off = 32; off > 0; off >>= 1) ws += static_cast<uint32_t>(__shfl_xor(static_cast<int>(ws), off, 64));
    if (lane == 0) {
        wa[slice] = (a + PACK - 1) / PACK * PACK;
        wb[slice] = (b + PACK - 1) / PACK * PACK;
        bsum[slice] = ws;
        atomicAdd(nnz, sum);
        atomicAdd(nnz + 1, sum_b);
    }
}

// Section-A level codes (one wave per slice) from the column pass (columns in
// the full packed layout, the group of each section-A entry level-major, -1 on
// padding): del = d when every row r of the slice has column r + d at that
// level (d = 0: a level of padding only, whose values are 0), else
// LEVEL_GATHER; code = -1 when every lane's value is the same by construction
// (the same group, rows that agree on every bit of the group's Z masks, and a
// group summed the same way for every row: a sign table, or too short for
// term-parallel sums; or padding only), else the level's index among the
// slice's levels that store 64 values.  n_stored / n_vals: the slice's
// LEVEL_GATHER levels / levels with 64 values; pad[0..3) (one atomic each per
// wave): padding lanes of the levels with 64 values, padding lanes of the
// LEVEL_GATHER levels, levels of padding only.
__global__ __launch_bounds__(256) void k_cache_levels(const int32_t *__restrict__ col_full,
                                                      const int32_t *__restrict__ gid_a,
                                                      const int64_t *__restrict__ off, const uint32_t *__restrict__ wa,
                                                      const int64_t *__restrict__ aoff, const uint64_t *__restrict__ keys,
                                                      const uint64_t *__restrict__ g_zor,
                                                      const int32_t *__restrict__ g_tab,
                                                      const uint32_t *__restrict__ g_toff, int64_t n_states,
                                                      int64_t n_slices, int32_t *__restrict__ adel,
                                                      int32_t *__restrict__ acode, uint32_t *__restrict__ n_stored,
                                                      uint32_t *__restrict__ n_vals, unsigned long long *pad) {
    const int64_t slice = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / SLICE;
    if (slice >= n_slices) return;
    const int lane = static_cast<int>(threadIdx.x % SLICE);
    const int64_t row = slice * SLICE + lane;
    const bool valid = row < n_states;
    const uint64_t s_me = valid ? keys[row] : 0ull;
    const uint64_t s_0 = __shfl(s_me, 0, 64);                 // lane 0 always holds a row
    const uint32_t w = wa[slice];
    const int64_t o = off[slice], ao = aoff[slice];
    uint32_t stored = 0, vals = 0;
    unsigned long long pad_vals = 0, pad_cols = 0, pad_levels = 0;
    for (uint32_t k = 0; k < w; ++k) {
        const int64_t slot = packed_slot<int32_t>(o, k, lane);
        const int32_t d = valid ? col_full[slot] - static_cast<int32_t>(row) : 0;
        const int32_t d0 = __shfl(d, 0, 64);
        const bool affine = __ballot(valid && d != d0) == 0;
        const int32_t g = valid ? gid_a[(ao + k) * SLICE + lane] : -1;
        const int32_t g0 = __shfl(g, 0, 64);
        bool uniform = __ballot(valid && g != g0) == 0;
        if (uniform && g0 >= 0) {
            const uint32_t len = g_toff[g0 + 1] - g_toff[g0];
            uniform = (g_tab[g0] >= 0 || len < static_cast<uint32_t>(TP_MIN_TERMS)) &&
                      __ballot(valid && ((s_me ^ s_0) & g_zor[g0]) != 0) == 0;
        }
        if (lane == 0) {
            adel[ao + k] = affine ? d0 : LEVEL_GATHER;
            acode[ao + k] = uniform ? -1 : static_cast<int32_t>(vals);
        }
        stored += affine ? 0u : 1u;
        vals += uniform ? 0u : 1u;
        const unsigned long long n_pad = __popcll(__ballot(!(valid && g >= 0)));
        pad_vals += uniform ? 0ull : n_pad;
        pad_cols += affine ? 0ull : n_pad;
        pad_levels += n_pad == SLICE ? 1ull : 0ull;
    }
    if (lane == 0) {
        n_stored[slice] = stored;
        n_vals[slice] = vals;
        atomicAdd(pad, pad_vals);
        atomicAdd(pad + 1, pad_cols);
        atomicAdd(pad + 2, pad_levels);
    }
}

// Stored columns (one wave per slice) from coff[s] on: the LEVEL_GATHER
// section-A levels (64 lane-contiguous columns each), then section B in SELL
// order (each row's packets up to its chunk width, from the full layout).
__global__ __launch_bounds__(256) void k_cache_columns(const int32_t *__restrict__ col_full,
                                                       const int64_t *__restrict__ off, const uint32_t *__restrict__ wa,
                                                       const int64_t *__restrict__ aoff,
                                                       const int32_t *__restrict__ adel, const uint8_t *__restrict__ bpos,
                                                       const uint16_t *__restrict__ bw,
                                                       const int64_t *__restrict__ coff, int64_t n_states,
                                                       int64_t n_slices, int32_t *__restrict__ col) {
    const int64_t slice = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / SLICE;
    if (slice >= n_slices) return;
    const int lane = static_cast<int>(threadIdx.x % SLICE);
    const int64_t row = slice * SLICE + lane;
    const bool valid = row < n_states;
    const uint32_t w_a = wa[slice];
    const int64_t o = off[slice], ao = aoff[slice];
    int64_t dst = coff[slice];
    for (uint32_t k = 0; k < w_a; ++k) {
        if (adel[ao + k] != LEVEL_GATHER) continue;
        col[dst + lane] = valid ? col_full[packed_slot<int32_t>(o, k, lane)] : 0;
        dst += SLICE;
    }
    uint32_t W[NCH];
    load_chunk_widths(bw, slice, W);
    const int pos = bpos[row];
    const uint32_t w_me = W[pos / CH];
    for (uint32_t q = 0; q < w_me; ++q) {
        i32x4 t = {0, 0, 0, 0};
        if (valid) t = *reinterpret_cast<const i32x4 *>(col_full + packed_slot<int32_t>(o, w_a + q * PACK, lane));
        *reinterpret_cast<i32x4 *>(col + sell_slot<int32_t>(dst, sell_level(W, q), 0, pos)) = t;
    }
}

// Full-space action (no subspace index): y[b, c] = sum_j C_j(b) x[b ^ x_j, c] over
// every basis state b of 2^n, for n_cols right-hand-side columns stored row major,
// C_j(b) = sum_k coef_k (-1)^popcount(b & z_k) summed term by term in group order.
// One thread per (b, c); the groups' masks and coefficients are wave-uniform
// scalar loads, and the 64 lanes of a wave gather from one aligned block of x
// (b ^ x_j permutes the low bits of a run of consecutive b).
template <typename V, typename R, bool REALC>
__global__ __launch_bounds__(256) void k_full_apply(const uint64_t *__restrict__ g_x,
                                                    const uint32_t *__restrict__ g_toff,
                                                    const uint64_t *__restrict__ z, const void *__restrict__ coef_v,
                                                    int32_t n_groups, int64_t n_items, int64_t n_cols,
                                                    const V *__restrict__ x, V *__restrict__ y) {
    using CC = typename C2<R>::T;
    const CC *coef = static_cast<const CC *>(coef_v);
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n_items;
         t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t col = t % n_cols;
        const uint64_t b = static_cast<uint64_t>(t / n_cols);
        V acc = vzero(static_cast<V *>(nullptr));
        for (int32_t j = 0; j < n_groups; ++j) {
            const uint64_t xj = ldc(g_x + j);
            const uint32_t k0 = ldc(g_toff + j), k1 = ldc(g_toff + j + 1);
            R cr = R(0), ci = R(0);
            for (uint32_t k = k0; k < k1; ++k) {
                const CC c = ldc(coef + k);
                const bool neg = __popcll(b & ldc(z + k)) & 1;
                cr += neg ? -c.x : c.x;
                if constexpr (!REALC) ci += neg ? -c.y : c.y;
            }
            const V xv = x[static_cast<int64_t>(b ^ xj) * n_cols + col];
            if constexpr (REALC) {
                fma_acc(acc, cr, xv);
            } else {
                CC cc;
                cc.x = cr;
                cc.y = ci;
                fma_acc(acc, cc, xv);
            }
        }
        y[t] = acc;
    }
}

__global__ void k_eigen_batch(const double *__restrict__ mats, int count, int n, double *__restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) eigen_small(mats + static_cast<int64_t>(k) * 2 * n * n, n, out + static_cast<int64_t>(k) * (2 + 2 * n));
}

// ---------------------------------------------------------------- host: index build
inline int highest_diff_bit(uint64_t a, uint64_t b) { return 63 - __builtin_clzll(a ^ b); }

inline int ceil_log2(uint64_t n) {
    int l = 0;
    while ((uint64_t(1) << l) < n) ++l;
    return l;
}

struct Segmentation {
    int L = 0;
    std::vector<uint64_t> seg_hi;
    std::vector<uint32_t> seg_off;
    std::vector<uint64_t> cls_hi;
    std::vector<uint32_t> cls_goff;
    int max_seg = 0;
    double cost = 0.0;
};

// maximum segment size for split L, from hd[i] = highest differing bit of keys i-1, i
int max_segment(const std::vector<int8_t> &hd, int L) {
    int best = hd.empty() ? 0 : 1, run = 1;
    for (size_t i = 1; i < hd.size(); ++i) {
        if (hd[i] >= L) {
            run = 1;
        } else {
            ++run;
            best = std::max(best, run);
        }
    }
    return best;
}

// the longest segment and the classes (runs of equal xmask >> L) of a segmentation
// whose seg_hi / seg_off are set
void segment_classes(Segmentation &sg, const uint64_t *xmask, int32_t J) {
    const int L = sg.L;
    sg.cls_hi.clear();
    sg.cls_goff.clear();
    sg.max_seg = 0;
    for (size_t m = 0; m + 1 < sg.seg_off.size(); ++m)
        sg.max_seg = std::max<int>(sg.max_seg, sg.seg_off[m + 1] - sg.seg_off[m]);
    for (int32_t j = 0; j < J; ++j) {
        const uint64_t h = xmask[j] >> L;
        if (j == 0 || h != sg.cls_hi.back()) {
            sg.cls_hi.push_back(h);
            sg.cls_goff.push_back(static_cast<uint32_t>(j));
        }
    }
    sg.cls_goff.push_back(static_cast<uint32_t>(J));
}

void build_segments(Segmentation &sg, const uint64_t *keys, int64_t n, const uint64_t *xmask, int32_t J) {
    const int L = sg.L;
    sg.seg_hi.clear();
    sg.seg_off.clear();
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t h = (L >= 64) ? 0 : keys[i] >> L;
        if (i == 0 || h != sg.seg_hi.back()) {
            sg.seg_hi.push_back(h);
            sg.seg_off.push_back(static_cast<uint32_t>(i));
        }
    }
    sg.seg_off.push_back(static_cast<uint32_t>(n));
    segment_classes(sg, xmask, J);
}

// What the cost model needs of a product subspace's alpha lists: the product
// split (0: none), the groups the lists would hold, and their entries per row
struct AlphaModel {
    int32_t a_bits = 0;
    std::vector<uint8_t> listed;
    double hits_per_row = 0.0;
};

// Estimated per-thread work of the tile kernel, from a deterministic sample of
// tiles.  Below the product split, an alpha class searches only its groups
// outside the lists (none: no staging either) and every row adds its list
// entries (a load and a multiply-add each).
double estimate_cost(const Segmentation &sg, bool has_diag, const AlphaModel &am) {
    const int M = static_cast<int>(sg.seg_hi.size());
    const int C = static_cast<int>(sg.cls_hi.size());
    std::vector<uint32_t> tile_seg;
    for (int m = 0; m < M; ++m) {
        const uint32_t sz = sg.seg_off[m + 1] - sg.seg_off[m];
        for (uint32_t r = 0; r < sz; r += TILE) tile_seg.push_back(m);
    }
    if (tile_seg.empty()) return 0.0;
    const size_t n_sample = std::min<size_t>(tile_seg.size(), 128);
    const int lm = ceil_log2(static_cast<uint64_t>(M));
    const bool lists = am.a_bits > sg.L && !am.listed.empty();
    std::vector<uint32_t> cls_listed(static_cast<size_t>(C), 0);   // listed groups of each alpha class
    if (lists)
        for (int c = 0; c < C && (sg.cls_hi[c] >> (am.a_bits - sg.L)) == 0; ++c)
            for (uint32_t j = sg.cls_goff[c]; j < sg.cls_goff[c + 1]; ++j) cls_listed[c] += am.listed[j];
    double total = 0.0;
    for (size_t si = 0; si < n_sample; ++si) {
        const uint32_t m = tile_seg[(si * tile_seg.size()) / n_sample];
        double cost = ((C + TILE - 1) / TILE) * (lm + 8.0);
        for (int c = 0; c < C; ++c) {
            const uint64_t t = sg.seg_hi[m] ^ sg.cls_hi[c];
            auto it = std::lower_bound(sg.seg_hi.begin(), sg.seg_hi.end(), t);
            if (it == sg.seg_hi.end() || *it != t) continue;
            const int ts = static_cast<int>(it - sg.seg_hi.begin());
            const uint32_t tn = sg.seg_off[ts + 1] - sg.seg_off[ts];
            uint32_t groups = sg.cls_goff[c + 1] - sg.cls_goff[c];
            if (has_diag && sg.cls_goff[c] == 0) groups -= 1;
            if (cls_listed[c] > 0) {
                groups -= cls_listed[c];
                if (groups == 0) continue;
            }
            cost += (tn + TILE - 1) / TILE + 12.0 + groups * (2.0 * ceil_log2(tn) + 4.0);
        }
        if (lists) cost += am.hits_per_row * 6.0;
        total += cost;
    }
    return total / n_sample * tile_seg.size();
}

}  // namespace

struct ProbHost {            // host record of one problem of a plan
    int64_t row_base = 0, n_states = 0;
    int64_t term_base = 0, n_terms = 0;
    int32_t grp_base = 0, n_groups = 0;
    int32_t seg_base = 0, n_segs = 0;
    int32_t cls_base = 0, n_classes = 0;
    int32_t n_tiles = 0, max_seg = 0, L = 0, n_qubits = 0, a_bits = 0;
    int32_t has_diag = 0, max_group_terms = 0;
    int32_t max_offdiag_terms = 0;   // longest group with x != 0
    uint32_t diag_terms = 0;   // terms of the x == 0 group
    int32_t is_real = 1;
    double cost = 0.0;         // cost-model estimate of one matvec
    int64_t a_rows = 0;        // low strings of the product at the section split (0: no product)
    uint32_t al_nlo = 0;       // alpha lists: low strings (0: none), first alpha index,
    uint32_t al_a0 = 0;        //   and the problem's listed groups [al_lg0, al_lg0 + al_lgn) of the plan's list
    uint32_t al_lg0 = 0, al_lgn = 0;
    uint32_t al_nhi = 0;       // high strings of the product
    uint32_t al_g0 = 0;        // its G tables' first element (al_rn values per beta block)
    uint32_t al_r0 = 0, al_rn = 0;   // the problem's G-table rows [al_r0, al_r0 + al_rn) of the plan's list
    uint32_t al_hsl = 0;       // log2 of the table values per alpha pattern the product kernel stages
    uint32_t al_pat0 = 0, al_npat = 0;   // its alpha patterns [al_pat0, al_pat0 + al_npat) of the plan's list,
    uint32_t al_bpat0 = 0, al_nbpat = 0; //   its beta patterns of the plan's list,
    uint32_t pm_base = UINT32_MAX;       //   its partner map's first element
    uint32_t bp_b0 = 0;                  //   and its first beta index of the beta-partner offsets
    uint32_t full = 0;                   // beta lists: every group listed (no class searched),
    uint32_t bl_b0 = 0;                  //   the first beta index of their offsets,
    uint32_t bl_lg0 = 0, bl_lgn = 0;     //   the problem's beta groups of the plan's list,
    uint32_t bl_r0 = 0, bl_rn = 0;       //   and its G-table rows (entries per alpha string)
};

struct HitCache {            // sliced-ELL cache of the hits (column, C_j(s_i)) in one precision
    void *block = nullptr;     // per-slice arrays, level codes, section-B and single values
    void *col_block = nullptr; // stored columns
    void *vala_block = nullptr;// section-A values of the levels with 64
    void *sell_block = nullptr;// section B's SELL order: bpos, bperm, bw
    uint8_t *bpos = nullptr;   // each row's position in its slice (descending section-B count)
    uint8_t *bperm = nullptr;  // the row (within the slice) at each position
    uint16_t *bw = nullptr;    // the 8 chunk widths (packets) of each slice
    int2 *win = nullptr;       // the cached matvec's staging window of each workgroup (first row, rows)
    int32_t win_max = 0;       // the widest window (rows)
    int32_t *col = nullptr;    // stored columns: gathered section-A levels and section B of each slice
    void *valb = nullptr;      // section-B values (packets)
    void *vala = nullptr;      // section-A values of the levels with 64 (lane-contiguous per level)
    void *aval = nullptr;      // the value of each section-A level with one value
    int64_t *boff = nullptr;   // first section-B value of each slice
    int64_t *avoff = nullptr;  // first vala value of each slice
    int64_t *coff = nullptr;   // first stored column of each slice
    int64_t *aoff = nullptr;   // first section-A level of each slice
    int32_t *adel = nullptr;   // section-A level codes: row offset d, or LEVEL_GATHER
    int32_t *acode = nullptr;  // section-A level codes: -1 (one value, aval) or index among the vala levels
    uint32_t *wa = nullptr;    // section-A width of each slice
    uint32_t *wb = nullptr;    // section-B width of each slice
    int64_t entries = 0, nnz = 0, n_slices = 0, bytes = 0;   // entries: stored values
    int64_t full_entries = 0, col_entries = 0, a_levels = 0, valb_entries = 0, vala_entries = 0;
    int64_t b_hits = 0;        // section-B hits (the rest of valb is padding)
    int64_t pad_vala = 0, pad_cola = 0, pad_levels = 0;   // padding of the section-A levels (k_cache_levels)
    int cplx = 0;
    bool built = false;
    bool shared = false;       // both precisions built together: the FP64 cache owns the allocations
};

// workgroups per problem of the batched vector kernels (256 threads each): about
// 2048 rows per workgroup, at most 1024
inline int vector_chunks(int64_t max_rows) {
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(1024, (max_rows + 2047) / 2048)));
}

struct rqx_plan {
    HitCache cache[2];         // [0]: FP32 values, [1]: FP64 values
    int use_cache = 2;         // 0: never, 1: a built cache always, 2: the faster path (cache_slower)
    int8_t cache_slower[4] = {0, 0, 0, 0};   // per dtype: the product kernel measured faster than the cache
    int device = 0;
    int is_real = 1;
    int long_groups = 0;       // some group has a sign table, or x != 0 and TP_MIN_TERMS or more terms
    int all_full = 0;          // every problem fully listed (the matvec runs k_product)
    int max_seg = 0;
    int persistent = 0;        // launch the work-queue kernel
    int blocks_per_cu = 0;     // persistent grid: blocks per CU (0 = occupancy limit)
    int occupancy = 1;         // resident tile workgroups per CU (FP64 complex matvec kernel)
    mutable std::vector<std::pair<const void *, int>> occ_cache;   // per tile-kernel instance
    int n_cu = 1;
    int lds_max = 65536;       // the device's LDS per workgroup
    int64_t n_states = 0, n_total = 0, n_terms = 0;
    int32_t n_groups = 0, n_classes = 0, n_segs = 0, n_tiles = 0;
    std::vector<ProbHost> probs;
    // device arrays (one allocation, carved)
    void *block = nullptr;
    uint64_t *keys = nullptr;
    uint32_t *lo = nullptr;
    uint64_t *seg_hi = nullptr;
    uint32_t *seg_off = nullptr;
    uint64_t *cls_hi = nullptr;
    uint32_t *cls_goff = nullptr;
    uint32_t *cls_lane = nullptr;   // per class: every group a lane evaluates alone (sign table or short)
    uint32_t *g_xlo = nullptr;
    uint32_t *g_toff = nullptr;
    uint64_t *g_x = nullptr;
    uint64_t *z = nullptr;
    double2 *coef64 = nullptr;
    float2 *coef32 = nullptr;
    double2 *diag64 = nullptr;
    float2 *diag32 = nullptr;
    uint32_t *tile_row0 = nullptr;
    uint32_t *tile_seg = nullptr;
    uint32_t *tile_prob = nullptr;
    ProbDesc *prob = nullptr;
    int64_t *prob_rows = nullptr;   // K + 1 row offsets
    unsigned int *queue = nullptr;
    int chunks = 1;                 // workgroups per problem of the vector kernels
    double *dot_part = nullptr;     // their per-workgroup partial dots: K x chunks x 2 LP_DOT
    int hashed = 0;                 // targets found by hash tables (else binary search)
    int hash_words = 0;             // largest segment hash table, u32 entries
    uint32_t *h_off = nullptr;      // per segment: first hash-table entry
    uint32_t *h_log = nullptr;      // per segment: log2 of its bucket count
    uint32_t *h_tab = nullptr;      // hash tables
    void *hash_block = nullptr;     // h_off, h_log, h_tab of a device-built index (own allocation)
    void *st_block = nullptr;       // per-segment target lists (own allocation)
    uint32_t *st_off = nullptr;     // M + 1 offsets into st_list
    TargetDesc *st_list = nullptr;  // (class, target segment) pairs with their ranges
    uint64_t *g_zor = nullptr;      // per group: union of its terms' Z masks
    int32_t *g_tab = nullptr;       // per group: its sign table, or -1
    GroupDesc *gdesc = nullptr;     // per group: X mask, first Z mask, term range, sign table
    uint32_t n_tab = 0;             // sign tables (entry u of table t at u * n_tab + t)
    double *tab64re = nullptr;      // sign tables, FP64 and FP32, real and imaginary parts
    double *tab64im = nullptr;
    float *tab32re = nullptr;
    float *tab32im = nullptr;
    // alpha lists (own allocations): per alpha index its entries, the listed
    // groups of every problem, per class the searched groups of an alpha class;
    // then the first entry of each block of 64 alpha indices and the entries
    void *al_block = nullptr;
    void *al_ent_block = nullptr;
    uint32_t *al_len = nullptr;
    uint32_t *al_lg = nullptr;
    uint32_t *al_pg_off = nullptr;
    uint32_t *al_pg = nullptr;
    int64_t *al_boff = nullptr;
    AlEnt<float> *al_e32 = nullptr;
    AlEnt<double> *al_e64 = nullptr;
    uint32_t *al_eg = nullptr;      // per entry: its group
    AlEntC<float> *al_c32 = nullptr;   // the entries for real coefficients
    AlEntC<double> *al_c64 = nullptr;
    int64_t al_slots = 0;           // entries per precision, padding included
    SplitDesc *al_split = nullptr;  // per group: split values (split groups)
    uint32_t *al_rg = nullptr;      // G-table rows: group, first element, alpha pattern
    uint32_t *al_roff = nullptr;
    uint64_t *al_rz = nullptr;
    double2 *al_G64 = nullptr;      // G tables, FP64 and FP32
    float2 *al_G32 = nullptr;
    int64_t al_gelems = 0;
    void *al_pblock = nullptr;      // partner lists and class maps (own allocations)
    void *al_pent_block = nullptr;
    uint32_t *al_plen = nullptr;
    int64_t *al_pboff = nullptr;
    uint4 *al_pent = nullptr;
    uint32_t *al_bpoff = nullptr;
    uint2 *al_bpent = nullptr;
    int32_t *al_pmap = nullptr;
    uint64_t *al_ap = nullptr;      // alpha patterns and their reference Z parts
    uint64_t *al_pz = nullptr;
    void *bl_block = nullptr;       // beta lists (own allocation): offsets, listed groups, G-table rows, entries
    uint32_t *bl_off = nullptr;
    uint32_t *bl_lg = nullptr;
    uint32_t *bl_rg = nullptr;
    uint32_t *bl_roff = nullptr;
    uint64_t *bl_rz = nullptr;
    AlEnt<float> *bl_e32 = nullptr;
    AlEnt<double> *bl_e64 = nullptr;
    void *bl_ent_block = nullptr;   // their entries (own allocation)
    int64_t al_pslots = 0;
};

namespace {

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

struct ProblemIn;
bool has_sign_table(const ProblemIn &q, int32_t j);

struct ProblemIn {
    int32_t n_qubits;
    const uint64_t *xmask;
    const int64_t *grp_off;
    int32_t n_groups;
    const uint64_t *zmask;
    const double *coef;
    int64_t n_terms;
    const uint64_t *states;
    int64_t n_states;
};

// group j of q gets a sign table: off-diagonal, at least TAB_MIN_TERMS terms,
// at most TAB_MAX_BITS X bits, and every term's Z mask equal to the first
// one's outside the X bits
bool has_sign_table(const ProblemIn &q, int32_t j) {
    const uint64_t x = q.xmask[j];
    const int64_t k0 = q.grp_off[j], k1 = q.grp_off[j + 1];
    if (x == 0 || k1 - k0 < TAB_MIN_TERMS || __builtin_popcountll(x) > TAB_MAX_BITS) return false;
    for (int64_t t = k0 + 1; t < k1; ++t)
        if ((q.zmask[t] ^ q.zmask[k0]) & ~x) return false;
    return true;
}

// The distinct alpha parts (z & amask, first-seen order) of group j's terms that
// hold beta Z bits (z & ~amask != 0) into P; false when there are more than SPLIT_MAX
bool split_patterns(const ProblemIn &q, int32_t j, uint64_t amask, std::vector<uint64_t> &P) {
    P.clear();
    for (int64_t t = q.grp_off[j]; t < q.grp_off[j + 1]; ++t) {
        const uint64_t z = q.zmask[t];
        if ((z & ~amask) == 0) continue;
        if (std::find(P.begin(), P.end(), z & amask) != P.end()) continue;
        if (P.size() == static_cast<size_t>(SPLIT_MAX)) return false;
        P.push_back(z & amask);
    }
    return true;
}

// The cost model's view of the alpha lists of q when its states are a product
// at bit a_bits with n_lo low strings (first: the first n_lo states, ascending
// low parts): the groups the plan build lists (flip only bits below a_bits, a
// lane evaluates them alone, Z masks below a_bits) and their hits per row,
// counted for up to 256 low strings spread over the first block.  Empty when
// the states are no product, nothing would be listed, or the plan would launch
// the plain kernels (no sign table, no long group), which do not read lists.
AlphaModel alpha_model(const ProblemIn &q, const uint64_t *first, int64_t n_lo, int32_t a_bits) {
    AlphaModel am;
    if (n_lo <= 0 || a_bits <= 0 || a_bits >= 64) return am;
    const uint64_t amask = (1ull << a_bits) - 1ull;
    std::vector<uint8_t> listed(static_cast<size_t>(q.n_groups), 0);
    std::vector<uint64_t> lx, pat;   // X masks of the listed groups
    bool tp = false;
    for (int32_t j = 0; j < q.n_groups; ++j) {
        const uint64_t x = q.xmask[j];
        if (x == 0) continue;
        const bool tab = has_sign_table(q, j);
        const int64_t k = q.grp_off[j + 1] - q.grp_off[j];
        tp = tp || tab || k >= TP_MIN_TERMS;
        if ((x & ~amask) != 0) continue;
        uint64_t zor = 0;
        for (int64_t t = q.grp_off[j]; t < q.grp_off[j + 1]; ++t) zor |= q.zmask[t];
        if (((tab || k < TP_MIN_TERMS) && (zor & ~amask) == 0) || split_patterns(q, j, amask, pat)) {
            listed[static_cast<size_t>(j)] = 1;
            lx.push_back(x);
        }
    }
    if (!tp || lx.empty()) return am;
    std::vector<uint64_t> al(static_cast<size_t>(n_lo));
    for (int64_t i = 0; i < n_lo; ++i) al[static_cast<size_t>(i)] = first[i] & amask;
    const int64_t ns = std::min<int64_t>(n_lo, 256);
    int64_t hits = 0;
    for (int64_t u = 0; u < ns; ++u) {
        const uint64_t a = al[static_cast<size_t>((u * (n_lo - 1)) / std::max<int64_t>(ns - 1, 1))];
        for (uint64_t x : lx) hits += std::binary_search(al.begin(), al.end(), a ^ x) ? 1 : 0;
    }
    am.a_bits = a_bits;
    am.listed.swap(listed);
    am.hits_per_row = static_cast<double>(hits) / static_cast<double>(ns);
    return am;
}

// two-choice cuckoo table of one segment's sorted low keys (2^lg buckets of two
// entries, key << HASH_RANK_BITS | rank); false if the insertions do not settle.
// One routine for the host build and the device build (one thread per
// segment), so both give the same table.
__host__ __device__ bool cuckoo_build(const uint32_t *keys, uint32_t n, uint32_t lg, uint32_t *tab) {
    const uint32_t nb = 1u << lg;
    for (uint32_t i = 0; i < 2 * nb; ++i) tab[i] = HASH_EMPTY;
    uint64_t rng = 0x9E3779B97F4A7C15ull ^ n;
    for (uint32_t r = 0; r < n; ++r) {
        uint32_t e = (keys[r] << HASH_RANK_BITS) | r;
        bool placed = false;
        for (int kick = 0; kick < 512 && !placed; ++kick) {
            const uint32_t k = e >> HASH_RANK_BITS;
            const uint32_t c[2] = {hash_b1(k, lg), hash_b2(k, lg)};
            for (int i = 0; i < 2 && !placed; ++i)
                for (int j = 0; j < 2 && !placed; ++j)
                    if (tab[2 * c[i] + j] == HASH_EMPTY) {
                        tab[2 * c[i] + j] = e;
                        placed = true;
                    }
            if (placed) break;
            rng = rng * 6364136223846793005ull + 1442695040888963407ull;
            const uint32_t slot = 2 * c[(rng >> 33) & 1] + static_cast<uint32_t>((rng >> 34) & 1);
            const uint32_t evicted = tab[slot];
            tab[slot] = e;
            e = evicted;
        }
        if (!placed) return false;
    }
    return true;
}

// Hash tables of every segment (lo: low keys, seg_off: M + 1 row offsets), built
// on the host threads; h_log[m] = log2 buckets, h_off[m] = first entry.  False if
// some segment's table does not settle (the plan then uses binary search).
// Returns false (the plan then uses binary search) when a segment's table cannot be
// built within `max_words` u32 entries, the tile kernel's LDS limit per workgroup.
bool build_hash_tables(const std::vector<uint32_t> &lo, const std::vector<uint32_t> &seg_off, size_t max_words,
                       std::vector<uint32_t> &h_off, std::vector<uint32_t> &h_log, std::vector<uint32_t> &h_tab) {
    const size_t M = seg_off.size() - 1;
    std::vector<std::vector<uint32_t>> tabs(M);
    h_log.assign(M, 0);
    std::vector<char> ok(M, 1);
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto work = [&](unsigned w) {
        for (size_t m = w; m < M; m += nt) {
            const uint32_t n = seg_off[m + 1] - seg_off[m];
            uint32_t lg = 1;
            while ((2u << lg) * 7u < n * 10u) ++lg;     // load <= 0.7
            bool done = false;
            for (int tries = 0; tries < 3 && !done && (size_t(2) << lg) <= max_words; ++tries, ++lg) {
                tabs[m].resize(size_t(2) << lg);
                done = cuckoo_build(lo.data() + seg_off[m], n, lg, tabs[m].data());
                if (done) h_log[m] = lg;
            }
            ok[m] = done;
        }
    };
    std::vector<std::thread> pool;
    for (unsigned w = 1; w < nt; ++w) pool.emplace_back(work, w);
    work(0);
    for (std::thread &t : pool) t.join();
    h_off.assign(M, 0);
    uint64_t total = 0;
    for (size_t m = 0; m < M; ++m) {
        if (!ok[m]) return false;
        h_off[m] = static_cast<uint32_t>(total);
        total += tabs[m].size();
        if (total > UINT32_MAX) return false;
    }
    h_tab.resize(static_cast<size_t>(total));
    for (size_t m = 0; m < M; ++m) std::copy(tabs[m].begin(), tabs[m].end(), h_tab.begin() + h_off[m]);
    return true;
}

// dynamic LDS of the tile kernel: one staged segment of uint32 low keys, or its hash table
size_t seg_lds_bytes(const rqx_plan *p) {
    return static_cast<size_t>(std::max(p->hashed ? p->hash_words : p->max_seg, 1)) * 4;
}

// (re)compute the sign tables from the plan's coefficients
void launch_group_tables(const rqx_plan *p, hipStream_t st) {
    if (p->n_groups == 0) return;
    const unsigned blocks = static_cast<unsigned>((p->n_groups + 255) / 256);
    hipLaunchKernelGGL(k_group_tables, dim3(blocks), dim3(256), 0, st, p->g_tab, p->g_x, p->g_toff, p->z, p->coef64,
                       p->coef32, p->n_groups, p->n_tab, p->tab64re, p->tab64im, p->tab32re, p->tab32im);
}

// per segment, the classes whose target segment exists and that segment
// (two passes of k_seg_targets: counts, then the pairs at their offsets)
hipError_t build_targets(rqx_plan *p, hipStream_t st) {
    const size_t M = static_cast<size_t>(std::max(p->n_segs, 0));
    uint32_t *cnt = nullptr;
    hipError_t e = hipMalloc(&cnt, std::max<size_t>(M, 1) * 4);
    if (e != hipSuccess) return e;
    e = fill_async(cnt, int64_t(0), static_cast<int64_t>(M), 0u, st);
    if (e != hipSuccess) {
        (void)hipFree(cnt);
        return e;
    }
    for (const ProbHost &ph : p->probs)
        if (ph.n_segs > 0 && ph.n_classes > 0)
            hipLaunchKernelGGL(k_seg_targets, dim3(ph.n_segs), dim3(TILE), 0, st, p->seg_hi, p->cls_hi,
                               static_cast<uint32_t>(ph.seg_base), static_cast<uint32_t>(ph.n_segs),
                               static_cast<uint32_t>(ph.cls_base), static_cast<uint32_t>(ph.n_classes),
                               cnt + ph.seg_base, nullptr, nullptr, nullptr, nullptr, nullptr, 0u, nullptr);
    std::vector<uint32_t> hc(M);
    e = hipGetLastError();
    if (e == hipSuccess && M) e = hipMemcpyAsync(hc.data(), cnt, M * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(cnt);
    if (e != hipSuccess) return e;
    std::vector<uint32_t> off(M + 1, 0);
    uint64_t total = 0;
    for (size_t m = 0; m < M; ++m) {
        off[m] = static_cast<uint32_t>(total);
        total += hc[m];
        if (total > UINT32_MAX) return hipErrorInvalidValue;
    }
    off[M] = static_cast<uint32_t>(total);
    const size_t o_list = align_up((M + 1) * 4);
    e = hipMalloc(&p->st_block, o_list + std::max<size_t>(total, 1) * sizeof(TargetDesc));
    if (e != hipSuccess) return e;
    p->st_off = static_cast<uint32_t *>(p->st_block);
    p->st_list = reinterpret_cast<TargetDesc *>(static_cast<char *>(p->st_block) + o_list);
    e = hipMemcpyAsync(p->st_off, off.data(), off.size() * 4, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    for (const ProbHost &ph : p->probs)
        if (ph.n_segs > 0 && ph.n_classes > 0)
            hipLaunchKernelGGL(k_seg_targets, dim3(ph.n_segs), dim3(TILE), 0, st, p->seg_hi, p->cls_hi,
                               static_cast<uint32_t>(ph.seg_base), static_cast<uint32_t>(ph.n_segs),
                               static_cast<uint32_t>(ph.cls_base), static_cast<uint32_t>(ph.n_classes), nullptr,
                               p->st_off, p->seg_off, p->hashed ? p->h_off : nullptr, p->h_log, p->cls_goff,
                               ph.has_diag ? static_cast<uint32_t>(ph.grp_base) : UINT32_MAX, p->st_list);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);   // off is released on return
    return e;
}

KParams make_params(const rqx_plan *p, bool fp64) {
    KParams k{};
    k.keys = p->keys;
    k.lo = p->lo;
    k.seg_hi = p->seg_hi;
    k.seg_off = p->seg_off;
    k.cls_hi = p->cls_hi;
    k.cls_goff = p->cls_goff;
    k.cls_lane = p->cls_lane;
    k.st_off = p->st_off;
    k.h_off = p->hashed ? p->h_off : nullptr;
    k.h_log = p->h_log;
    k.h_tab = p->h_tab;
    k.st_list = p->st_list;
    k.g_xlo = p->g_xlo;
    k.g_toff = p->g_toff;
    k.g_x = p->g_x;
    k.z = p->z;
    k.coef = fp64 ? static_cast<const void *>(p->coef64) : static_cast<const void *>(p->coef32);
    k.g_tab = p->g_tab;
    k.gdesc = p->gdesc;
    k.n_tab = p->n_tab;
    k.tab_re = fp64 ? static_cast<const void *>(p->tab64re) : static_cast<const void *>(p->tab32re);
    k.tab_im = fp64 ? static_cast<const void *>(p->tab64im) : static_cast<const void *>(p->tab32im);
    k.tab_lo_re = p->tab32re;
    k.tab_lo_im = p->tab32im;
    k.diag = fp64 ? static_cast<const void *>(p->diag64) : static_cast<const void *>(p->diag32);
    k.tile_row0 = p->tile_row0;
    k.tile_seg = p->tile_seg;
    k.tile_prob = p->tile_prob;
    k.prob = p->prob;
    k.queue = p->queue;
    k.n_total = p->n_total;
    k.n_tiles = static_cast<uint32_t>(p->n_tiles);
    k.al_len = p->al_len;
    k.al_boff = p->al_boff;
    k.al_ent = fp64 ? static_cast<const void *>(p->al_e64) : static_cast<const void *>(p->al_e32);
    k.al_pg_off = p->al_pg_off;
    k.al_pg = p->al_pg;
    k.al_ent_lo = p->al_e32;
    k.al_eg = p->al_eg;
    k.al_ec = fp64 ? static_cast<const void *>(p->al_c64) : static_cast<const void *>(p->al_c32);
    k.al_split = p->al_split;
    k.al_G = fp64 ? static_cast<const void *>(p->al_G64) : static_cast<const void *>(p->al_G32);
    k.al_G_lo = p->al_G32;
    k.al_plen = p->al_plen;
    k.al_pboff = p->al_pboff;
    k.al_pent = p->al_pent;
    k.al_bpoff = p->al_bpoff;
    k.al_bpent = p->al_bpent;
    k.al_pmap = p->al_pmap;
    k.al_ap = p->al_ap;
    k.al_pz = p->al_pz;
    k.bl_off = p->bl_off;
    k.bl_ent = fp64 ? static_cast<const void *>(p->bl_e64) : static_cast<const void *>(p->bl_e32);
    k.bl_ent_lo = p->bl_e32;
    return k;
}

// the alpha lists' count pass (fill = false: entries per alpha index) or fill
// pass (entries and values; also after a coefficient update)
hipError_t launch_alpha_lists(const rqx_plan *p, bool fill, hipStream_t st) {
    const KParams p64 = make_params(p, true), p32 = make_params(p, false);
    for (const ProbHost &ph : p->probs) {
        if (ph.al_nlo == 0) continue;
        // the G tables of the split groups (fill pass), at most 65535 rows a launch
        for (uint32_t r = 0; fill && r < ph.al_rn; r += 65535u) {
            const uint32_t nr = std::min(ph.al_rn - r, 65535u);
            hipLaunchKernelGGL(k_split_g, dim3((ph.al_nhi + 255) / 256, nr), dim3(256), 0, st, p64, p32,
                               p->keys + ph.row_base, ph.al_nlo, ph.al_nhi, ph.al_rn, ~((1ull << ph.a_bits) - 1ull),
                               p->al_rg + ph.al_r0 + r, p->al_rz + ph.al_r0 + r, p->al_roff + ph.al_r0 + r, p->al_G64,
                               p->al_G32);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k_alpha_lists, dim3(ph.al_nlo), dim3(256), 0, st, p64, p32, p->keys + ph.row_base,
                           ph.al_nlo, (1ull << ph.a_bits) - 1ull, p->al_lg + ph.al_lg0, ph.al_lgn, ph.al_a0,
                           p->al_len, fill ? p->al_boff : nullptr, p->al_e32, p->al_e64, p->al_eg);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (fill && p->al_slots > 0) {
        hipLaunchKernelGGL(k_compact_alpha, dim3(static_cast<uint32_t>((p->al_slots + 255) / 256)), dim3(256), 0, st,
                           p->al_e32, p->al_e64, p->al_c32, p->al_c64, p->al_slots);
        return hipGetLastError();
    }
    return hipSuccess;
}

// at most this many bytes of alpha-list entries (both precisions) per plan; a
// plan beyond it searches every group
constexpr int64_t ALPHA_LIST_MAX_BYTES = int64_t(1) << 30;

// Build the alpha lists of the problems that have them (ph.al_nlo > 0): counts,
// the block layout on the host, then the entries.  Without room for them the
// problems' descriptors are rewritten without lists.
hipError_t build_alpha_lists(rqx_plan *p, uint32_t n_alpha, const std::vector<uint32_t> &lg,
                             const std::vector<uint32_t> &pg_off, const std::vector<uint32_t> &pg,
                             const std::vector<SplitDesc> &split, const std::vector<uint32_t> &rg,
                             const std::vector<uint64_t> &rz, const std::vector<uint32_t> &roff, int64_t gelems,
                             std::vector<ProbDesc> &desc, hipStream_t st) {
    if (n_alpha == 0 || lg.empty()) return hipSuccess;
    // every problem searches all its groups: no lists built, or (plain kernels
    // only: no sign table and no long group in the plan) none read
    auto none = [&]() {
        for (size_t k = 0; k < p->probs.size(); ++k) {
            p->probs[k].al_nlo = 0;
            p->probs[k].full = 0;
            desc[k].al_nlo = 0;
            desc[k].full = 0;
        }
        hipError_t e = hipMemcpyAsync(p->prob, desc.data(), desc.size() * sizeof(ProbDesc), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);   // desc is the caller's host array
        return e;
    };
    if (!p->long_groups) return none();
    const size_t nr = rg.size();
    const size_t o_len = 0, o_lg = align_up(size_t(n_alpha) * 4), o_pgo = o_lg + align_up(lg.size() * 4),
                 o_pg = o_pgo + align_up(pg_off.size() * 4), o_split = o_pg + align_up(std::max<size_t>(pg.size(), 1) * 4),
                 o_rg = o_split + align_up(split.size() * sizeof(SplitDesc)), o_roff = o_rg + align_up(std::max<size_t>(nr, 1) * 4),
                 o_rz = o_roff + align_up(std::max<size_t>(nr, 1) * 4), bytes = o_rz + align_up(std::max<size_t>(nr, 1) * 8);
    hipError_t e = hipMalloc(&p->al_block, bytes);
    if (e != hipSuccess) return e;
    char *b = static_cast<char *>(p->al_block);
    p->al_len = reinterpret_cast<uint32_t *>(b + o_len);
    p->al_lg = reinterpret_cast<uint32_t *>(b + o_lg);
    p->al_pg_off = reinterpret_cast<uint32_t *>(b + o_pgo);
    p->al_pg = reinterpret_cast<uint32_t *>(b + o_pg);
    p->al_split = reinterpret_cast<SplitDesc *>(b + o_split);
    p->al_rg = reinterpret_cast<uint32_t *>(b + o_rg);
    p->al_roff = reinterpret_cast<uint32_t *>(b + o_roff);
    p->al_rz = reinterpret_cast<uint64_t *>(b + o_rz);
    e = fill_async(p->al_len, int64_t(0), static_cast<int64_t>(n_alpha), 0u, st);
    if (e == hipSuccess) e = hipMemcpyAsync(p->al_lg, lg.data(), lg.size() * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(p->al_pg_off, pg_off.data(), pg_off.size() * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && !pg.empty()) e = hipMemcpyAsync(p->al_pg, pg.data(), pg.size() * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && !split.empty())
        e = hipMemcpyAsync(p->al_split, split.data(), split.size() * sizeof(SplitDesc), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nr) e = hipMemcpyAsync(p->al_rg, rg.data(), nr * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nr) e = hipMemcpyAsync(p->al_roff, roff.data(), nr * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nr) e = hipMemcpyAsync(p->al_rz, rz.data(), nr * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_alpha_lists(p, false, st);
    std::vector<uint32_t> len(n_alpha);
    if (e == hipSuccess) e = hipMemcpyAsync(len.data(), p->al_len, size_t(n_alpha) * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    const uint32_t n_blk = n_alpha / 64;          // each problem's range is padded to a multiple of 64
    std::vector<int64_t> boff(n_blk);
    int64_t slots = 0;
    for (uint32_t k = 0; k < n_blk; ++k) {
        uint32_t w = 0;
        for (uint32_t u = 0; u < 64; ++u) w = std::max(w, len[64 * k + u]);
        boff[k] = slots;
        slots += 64 * static_cast<int64_t>(w);
    }
    const int64_t ent_bytes = slots * static_cast<int64_t>(sizeof(AlEnt<float>) + sizeof(AlEnt<double>) + 4 +
                                                           sizeof(AlEntC<float>) + sizeof(AlEntC<double>)) +
                              gelems * static_cast<int64_t>(sizeof(double2) + sizeof(float2));
    if (slots == 0 || ent_bytes > ALPHA_LIST_MAX_BYTES) return none();   // no hits at all, or no room
    const size_t o_e32 = align_up(size_t(n_blk) * 8), o_e64 = o_e32 + align_up(size_t(slots) * sizeof(AlEnt<float>)),
                 o_eg = o_e64 + align_up(size_t(slots) * sizeof(AlEnt<double>)),
                 o_c32 = o_eg + align_up(size_t(slots) * 4), o_c64 = o_c32 + align_up(size_t(slots) * sizeof(AlEntC<float>)),
                 o_g64 = o_c64 + align_up(size_t(slots) * sizeof(AlEntC<double>)),
                 o_g32 = o_g64 + align_up(size_t(std::max<int64_t>(gelems, 1)) * sizeof(double2));
    e = hipMalloc(&p->al_ent_block, o_g32 + size_t(std::max<int64_t>(gelems, 1)) * sizeof(float2));
    if (e != hipSuccess) return e;
    char *eb = static_cast<char *>(p->al_ent_block);
    p->al_boff = reinterpret_cast<int64_t *>(eb);
    p->al_e32 = reinterpret_cast<AlEnt<float> *>(eb + o_e32);
    p->al_e64 = reinterpret_cast<AlEnt<double> *>(eb + o_e64);
    p->al_eg = reinterpret_cast<uint32_t *>(eb + o_eg);
    p->al_c32 = reinterpret_cast<AlEntC<float> *>(eb + o_c32);
    p->al_c64 = reinterpret_cast<AlEntC<double> *>(eb + o_c64);
    p->al_G64 = reinterpret_cast<double2 *>(eb + o_g64);
    p->al_G32 = reinterpret_cast<float2 *>(eb + o_g32);
    p->al_slots = slots;
    p->al_gelems = gelems;
    e = hipMemcpyAsync(p->al_boff, boff.data(), boff.size() * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_alpha_lists(p, true, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);   // boff is released on return
    return e;
}

// at most this many class-map elements (int32) per plan
constexpr int64_t CMAP_MAX_ELEMS = int64_t(1) << 24;

// Partner lists of the problems with a partner map (ph.pm_base set): alpha
// partners in the blocked layout of the alpha lists, beta partners in offsets per
// beta index; then the map.  Without them (no alpha lists, no room) the problems'
// descriptors are rewritten without a map and their classes are searched.
hipError_t build_partners(rqx_plan *p, uint32_t n_alpha, uint32_t n_boff, const std::vector<uint64_t> &apats,
                          const std::vector<uint64_t> &apz, const std::vector<uint64_t> &bpats,
                          const std::vector<int32_t> &pmap, std::vector<ProbDesc> &desc, hipStream_t st) {
    if (pmap.empty()) return hipSuccess;
    auto none = [&]() {
        for (size_t k = 0; k < p->probs.size(); ++k) {
            p->probs[k].pm_base = UINT32_MAX;
            p->probs[k].full = 0;
            desc[k].pm_base = UINT32_MAX;
            desc[k].full = 0;
        }
        hipError_t e = hipMemcpyAsync(p->prob, desc.data(), desc.size() * sizeof(ProbDesc), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);   // desc is the caller's host array
        return e;
    };
    if (!p->al_ent_block || n_alpha == 0 || apats.empty() || bpats.empty()) return none();
    const size_t o_len = 0, o_ap = align_up(size_t(n_alpha) * 4), o_pz = o_ap + align_up(apats.size() * 8),
                 o_bp = o_pz + align_up(apz.size() * 8), o_map = o_bp + align_up(bpats.size() * 8),
                 o_boff = o_map + align_up(pmap.size() * 4),
                 bytes = o_boff + align_up(size_t(n_boff) * 4);
    hipError_t e = hipMalloc(&p->al_pblock, bytes);
    if (e != hipSuccess) return e;
    char *b = static_cast<char *>(p->al_pblock);
    uint32_t *plen = reinterpret_cast<uint32_t *>(b + o_len), *bpoff = reinterpret_cast<uint32_t *>(b + o_boff);
    uint64_t *ap = reinterpret_cast<uint64_t *>(b + o_ap), *pz = reinterpret_cast<uint64_t *>(b + o_pz),
             *bp = reinterpret_cast<uint64_t *>(b + o_bp);
    int32_t *map = reinterpret_cast<int32_t *>(b + o_map);
    e = fill_async(plen, int64_t(0), static_cast<int64_t>(n_alpha), 0u, st);
    if (e == hipSuccess) e = fill_async(bpoff, int64_t(0), static_cast<int64_t>(n_boff), 0u, st);
    if (e == hipSuccess) e = hipMemcpyAsync(ap, apats.data(), apats.size() * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(pz, apz.data(), apz.size() * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(bp, bpats.data(), bpats.size() * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(map, pmap.data(), pmap.size() * 4, hipMemcpyHostToDevice, st);
    // count (boff NULL) or fill passes of both partner kinds
    auto launch = [&](const int64_t *aboff, uint4 *aent, const uint32_t *boff, uint2 *bent) -> hipError_t {
        for (const ProbHost &ph : p->probs) {
            if (ph.pm_base == UINT32_MAX) continue;
            const uint64_t amask = (1ull << ph.a_bits) - 1ull;
            hipLaunchKernelGGL(k_alpha_partners, dim3(ph.al_nlo), dim3(256), 0, st, p->keys + ph.row_base, ph.al_nlo,
                               amask, ap + ph.al_pat0, pz + ph.al_pat0, ph.al_npat, ph.al_a0, plen, aboff, aent);
            hipError_t le = hipGetLastError();
            if (le != hipSuccess) return le;
            hipLaunchKernelGGL(k_beta_partners, dim3(ph.al_nhi), dim3(256), 0, st, p->keys + ph.row_base, ph.al_nlo,
                               ph.al_nhi, amask, bp + ph.al_bpat0, ph.al_nbpat, bpoff + ph.bp_b0,
                               boff ? boff + ph.bp_b0 : nullptr, bent);
            le = hipGetLastError();
            if (le != hipSuccess) return le;
        }
        return hipSuccess;
    };
    if (e == hipSuccess) e = launch(nullptr, nullptr, nullptr, nullptr);
    std::vector<uint32_t> len(n_alpha), bcnt(n_boff);
    if (e == hipSuccess) e = hipMemcpyAsync(len.data(), plen, size_t(n_alpha) * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(bcnt.data(), bpoff, size_t(n_boff) * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    const uint32_t n_blk = n_alpha / 64;
    std::vector<int64_t> aoff(n_blk);
    int64_t slots = 0;
    for (uint32_t k = 0; k < n_blk; ++k) {
        uint32_t w = 0;
        for (uint32_t u = 0; u < 64; ++u) w = std::max(w, len[64 * k + u]);
        aoff[k] = slots;
        slots += 64 * static_cast<int64_t>(w);
    }
    // beta offsets: each problem's nhi counts become nhi + 1 offsets (exclusive scan)
    std::vector<uint32_t> boffs(n_boff, 0);
    int64_t bents = 0;
    for (const ProbHost &ph : p->probs) {
        if (ph.pm_base == UINT32_MAX) continue;
        for (uint32_t j = 0; j <= ph.al_nhi; ++j) {
            boffs[ph.bp_b0 + j] = static_cast<uint32_t>(bents);
            if (j < ph.al_nhi) bents += bcnt[ph.bp_b0 + j];
        }
    }
    if (slots * 16 + bents * 8 > ALPHA_LIST_MAX_BYTES || bents > 0xFFFFFFFFLL) return none();   // no room
    const size_t o_aent = align_up(size_t(n_blk) * 8), o_bent = o_aent + align_up(size_t(std::max<int64_t>(slots, 1)) * 16);
    e = hipMalloc(&p->al_pent_block, o_bent + size_t(std::max<int64_t>(bents, 1)) * 8);
    if (e != hipSuccess) return e;
    char *eb = static_cast<char *>(p->al_pent_block);
    int64_t *pboff = reinterpret_cast<int64_t *>(eb);
    uint4 *pent = reinterpret_cast<uint4 *>(eb + o_aent);
    uint2 *bpent = reinterpret_cast<uint2 *>(eb + o_bent);
    e = hipMemcpyAsync(pboff, aoff.data(), aoff.size() * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(bpoff, boffs.data(), boffs.size() * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch(pboff, pent, bpoff, bpent);
    if (e == hipSuccess) e = hipStreamSynchronize(st);    // the host arrays are released on return
    if (e != hipSuccess) return e;
    p->al_plen = plen;
    p->al_pboff = pboff;
    p->al_pent = pent;
    p->al_bpoff = bpoff;
    p->al_bpent = bpent;
    p->al_pmap = map;
    p->al_ap = ap;
    p->al_pz = pz;
    p->al_pslots = slots;
    return hipSuccess;
}

// the beta lists' G-table rows and entries (fill = true; also after a
// coefficient update), or their counts per beta index into bl_off (fill = false)
hipError_t launch_beta_lists(const rqx_plan *p, bool fill, hipStream_t st) {
    const KParams p64 = make_params(p, true), p32 = make_params(p, false);
    for (const ProbHost &ph : p->probs) {
        if (!ph.full) continue;
        const uint64_t amask = (1ull << ph.a_bits) - 1ull;
        for (uint32_t r = 0; fill && r < ph.bl_rn; r += 65535u) {
            const uint32_t nr = std::min(ph.bl_rn - r, 65535u);
            hipLaunchKernelGGL(k_split_g, dim3((ph.al_nlo + 255) / 256, nr), dim3(256), 0, st, p64, p32,
                               p->keys + ph.row_base, 1u, ph.al_nlo, 1u, amask, p->bl_rg + ph.bl_r0 + r,
                               p->bl_rz + ph.bl_r0 + r, p->bl_roff + ph.bl_r0 + r, p->al_G64, p->al_G32);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k_beta_lists, dim3(ph.al_nhi), dim3(256), 0, st, p64, p32, p->keys + ph.row_base, ph.al_nlo,
                           ph.al_nhi, amask, p->bl_lg + ph.bl_lg0, ph.bl_lgn, p->bl_off + ph.bl_b0,
                           fill ? p->bl_off + ph.bl_b0 : nullptr, p->bl_e32, p->bl_e64);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Beta lists of the problems in the fully listed mode (ph.full): counts per beta
// index, their offsets on the host, the G rows and the entries.  Without them the
// problems' descriptors are rewritten out of that mode (their classes searched).
hipError_t build_beta_lists(rqx_plan *p, uint32_t n_boff, const std::vector<uint32_t> &lg,
                            const std::vector<uint32_t> &rg, const std::vector<uint64_t> &rz,
                            const std::vector<uint32_t> &roff, std::vector<ProbDesc> &desc, hipStream_t st) {
    if (lg.empty()) return hipSuccess;
    bool any = false;
    for (const ProbHost &ph : p->probs) any = any || ph.full;
    if (!any) return hipSuccess;
    auto none = [&]() {
        for (size_t k = 0; k < p->probs.size(); ++k) {
            p->probs[k].full = 0;
            desc[k].full = 0;
        }
        hipError_t e = hipMemcpyAsync(p->prob, desc.data(), desc.size() * sizeof(ProbDesc), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);   // desc is the caller's host array
        return e;
    };
    if (!p->al_ent_block) return none();
    const size_t nr = rg.size();
    const size_t o_off = 0, o_lg = align_up(size_t(n_boff) * 4), o_rg = o_lg + align_up(lg.size() * 4),
                 o_roff = o_rg + align_up(std::max<size_t>(nr, 1) * 4), o_rz = o_roff + align_up(std::max<size_t>(nr, 1) * 4),
                 bytes = o_rz + align_up(std::max<size_t>(nr, 1) * 8);
    hipError_t e = hipMalloc(&p->bl_block, bytes);
    if (e != hipSuccess) return e;
    char *b = static_cast<char *>(p->bl_block);
    p->bl_off = reinterpret_cast<uint32_t *>(b + o_off);
    p->bl_lg = reinterpret_cast<uint32_t *>(b + o_lg);
    p->bl_rg = reinterpret_cast<uint32_t *>(b + o_rg);
    p->bl_roff = reinterpret_cast<uint32_t *>(b + o_roff);
    p->bl_rz = reinterpret_cast<uint64_t *>(b + o_rz);
    e = fill_async(p->bl_off, int64_t(0), static_cast<int64_t>(n_boff), 0u, st);
    if (e == hipSuccess) e = hipMemcpyAsync(p->bl_lg, lg.data(), lg.size() * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nr) e = hipMemcpyAsync(p->bl_rg, rg.data(), nr * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nr) e = hipMemcpyAsync(p->bl_roff, roff.data(), nr * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nr) e = hipMemcpyAsync(p->bl_rz, rz.data(), nr * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_beta_lists(p, false, st);
    std::vector<uint32_t> cnt(n_boff);
    if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), p->bl_off, size_t(n_boff) * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    std::vector<uint32_t> off(n_boff, 0);
    int64_t total = 0;
    for (const ProbHost &ph : p->probs) {
        if (!ph.full) continue;
        for (uint32_t j = 0; j <= ph.al_nhi; ++j) {
            off[ph.bl_b0 + j] = static_cast<uint32_t>(total);
            if (j < ph.al_nhi) total += cnt[ph.bl_b0 + j];
        }
    }
    if (total * static_cast<int64_t>(sizeof(AlEnt<float>) + sizeof(AlEnt<double>)) > ALPHA_LIST_MAX_BYTES ||
        total > 0xFFFFFFFFLL)
        return none();
    const size_t o_e64 = align_up(size_t(std::max<int64_t>(total, 1)) * sizeof(AlEnt<float>));
    e = hipMalloc(&p->bl_ent_block, o_e64 + size_t(std::max<int64_t>(total, 1)) * sizeof(AlEnt<double>));
    if (e != hipSuccess) return e;
    p->bl_e32 = static_cast<AlEnt<float> *>(p->bl_ent_block);
    p->bl_e64 = reinterpret_cast<AlEnt<double> *>(static_cast<char *>(p->bl_ent_block) + o_e64);
    e = hipMemcpyAsync(p->bl_off, off.data(), off.size() * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_beta_lists(p, true, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);    // off is released on return
    return e;
}

int check_plan(const rqx_plan *p) {
    if (!p) return fail(RQX_ERR_ARG, "plan is NULL");
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != p->device)
        return fail(RQX_ERR_ARG, "plan was created on device " + std::to_string(p->device) +
                                     " but the current device is " + std::to_string(dev));
    return RQX_OK;
}

template <int MODE, typename R, bool CPLX, bool REALC = false, bool TP = true>
hipError_t launch_tiles(const rqx_plan *p, const KParams &k_in, const void *x, void *y, hipStream_t st) {
    if (p->n_tiles == 0) return hipSuccess;
    KParams k = k_in;
    size_t lds = seg_lds_bytes(p);
    if (MODE == MODE_MATVEC && TP) {   // the matvec's hit queues after the staged target
        k.hq_off = static_cast<uint32_t>((lds + 15) & ~size_t(15));
        lds = k.hq_off + (TILE / 64) * HQLayout<R, CPLX, REALC>::bytes;
    }
    if (p->persistent) {
        hipLaunchKernelGGL(k_fill<unsigned int>, dim3(1), dim3(64), 0, st, p->queue, int64_t(0), int64_t(1), 0u);
        int bpc = p->blocks_per_cu;
        if (bpc <= 0) {   // the occupancy of this very kernel (their register counts differ)
            const void *fn = reinterpret_cast<const void *>(&k_tiles_queue<MODE, R, CPLX, REALC, TP>);
            bpc = 0;
            for (const auto &e : p->occ_cache)
                if (e.first == fn) bpc = e.second;
            if (bpc <= 0) {
                int occ = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, TILE, lds) != hipSuccess || occ < 1)
                    occ = p->occupancy;
                p->occ_cache.emplace_back(fn, occ);
                bpc = occ;
            }
        }
        const int grid = std::max(1, std::min(p->n_tiles, p->n_cu * bpc));
        hipLaunchKernelGGL((k_tiles_queue<MODE, R, CPLX, REALC, TP>), dim3(grid), dim3(TILE), lds, st, k, x, y);
    } else {
        hipLaunchKernelGGL((k_tiles<MODE, R, CPLX, REALC, TP>), dim3(p->n_tiles), dim3(TILE), lds, st, k, x, y);
    }
    return hipGetLastError();
}

// The product kernel's LDS per workgroup for one problem (vector and coefficient
// element bytes): the beta block's x, its G values, the partner values per alpha
// pattern and their kinds
size_t prod_lds_bytes(const ProbHost &ph, size_t v_bytes, size_t c_bytes) {
    size_t b = ((static_cast<size_t>(ph.al_nlo) * v_bytes + 15) & ~size_t(15)) +
               ((static_cast<size_t>(ph.al_rn) * c_bytes + 15) & ~size_t(15));
    if (ph.pm_base != UINT32_MAX) b += static_cast<size_t>(ph.al_npat) * ((size_t(1) << ph.al_hsl) * c_bytes + 4);
    return b;
}

// whether the matvec runs the product kernel: every problem fully listed and its
// beta block within the device's LDS per workgroup (and the plan not on the
// plain kernels).  Blocks within PROD_LDS_BYTES leave room for two 1024-thread
// workgroups per CU; larger ones (complex128 blocks of 5000 rows) run one.
bool product_plan(const rqx_plan *p, size_t v_bytes, size_t c_bytes) {
    if (!p->long_groups || !p->all_full || p->probs.empty()) return false;
    const size_t budget = static_cast<size_t>(std::max(p->lds_max, PROD_LDS_BYTES));
    for (const ProbHost &ph : p->probs)
        if (prod_lds_bytes(ph, v_bytes, c_bytes) > budget) return false;
    return true;
}
bool is_fp64(int32_t dtype);
bool product_plan(const rqx_plan *p, int32_t dtype) {
    const size_t c = is_fp64(dtype) ? 16 : 8;
    const size_t v = dtype == RQX_C128 ? 16 : (dtype == RQX_F64 || dtype == RQX_C64) ? 8 : 4;
    return product_plan(p, v, c);
}

// matvec kernels with the plain evaluator where no group has a sign table or
// is long enough for term-parallel sums (other modes always carry the full one)
template <int MODE, typename R, bool CPLX, bool REALC>
hipError_t launch_tp(const rqx_plan *p, const KParams &k, const void *x, void *y, hipStream_t st) {
    if (MODE == MODE_MATVEC && !p->long_groups) return launch_tiles<MODE, R, CPLX, REALC, MODE != MODE_MATVEC>(p, k, x, y, st);
    if constexpr (MODE == MODE_MATVEC) {
        // every problem fully listed and its beta block (x, G values, the partner
        // values per alpha pattern) within the staging budget: the product kernel,
        // one launch per problem
        using V = typename VT<R, CPLX>::T;
        using CC = typename C2<R>::T;
        auto lds_bytes = [](const ProbHost &ph) { return prod_lds_bytes(ph, sizeof(V), sizeof(CC)); };
        if (product_plan(p, sizeof(V), sizeof(CC))) {
            for (size_t j = 0; j < p->probs.size(); ++j) {
                const ProbHost &ph = p->probs[j];
                const bool big = ph.al_nlo >= 2048;
                const uint32_t wgs = big ? 1024u : 256u, chunks = (ph.al_nlo + wgs - 1) / wgs;
                const size_t lds = lds_bytes(ph);
                if (big)
                    hipLaunchKernelGGL((k_product<R, CPLX, REALC, 1024>), dim3(ph.al_nhi * chunks), dim3(1024), lds, st, k,
                                       static_cast<uint32_t>(j), chunks, x, y);
                else
                    hipLaunchKernelGGL((k_product<R, CPLX, REALC, 256>), dim3(ph.al_nhi * chunks), dim3(256), lds, st, k,
                                       static_cast<uint32_t>(j), chunks, x, y);
                const hipError_t e = hipGetLastError();
                if (e != hipSuccess) return e;
            }
            return hipSuccess;
        }
    }
    return launch_tiles<MODE, R, CPLX, REALC>(p, k, x, y, st);
}

template <int MODE>
hipError_t launch_dtype(const rqx_plan *p, const KParams &k, const void *x, void *y, int32_t dtype, hipStream_t st) {
    constexpr bool MV = MODE == MODE_MATVEC;
    switch (dtype) {
        // real vectors imply real coefficients; complex vectors pick the kernel by the coefficients
        case RQX_F64: return launch_tp<MODE, double, false, MV>(p, k, x, y, st);
        case RQX_F32: return launch_tp<MODE, float, false, MV>(p, k, x, y, st);
        case RQX_C128:
            if (MV && p->is_real) return launch_tp<MODE, double, true, MV>(p, k, x, y, st);
            return launch_tp<MODE, double, true, false>(p, k, x, y, st);
        default:
            if (MV && p->is_real) return launch_tp<MODE, float, true, MV>(p, k, x, y, st);
            return launch_tp<MODE, float, true, false>(p, k, x, y, st);
    }
}

// host threads for the O(N) passes of a large subspace's index build (1 below 2^18 rows)
unsigned index_threads(int64_t n) {
    return n >= (int64_t(1) << 18) ? std::max(1u, std::min(16u, std::thread::hardware_concurrency())) : 1u;
}

// run f(w, begin, end) over nt contiguous chunks w of [0, n), one host thread each
template <typename F> void parallel_chunks(int64_t n, unsigned nt, F f) {
    if (nt <= 1) {
        f(0u, int64_t(0), n);
        return;
    }
    std::vector<std::thread> pool;
    for (unsigned w = 1; w < nt; ++w) pool.emplace_back(f, w, n * w / nt, n * (w + 1) / nt);
    f(0u, int64_t(0), n / nt);
    for (std::thread &t : pool) t.join();
}

int dev_check_states(const uint64_t *st, int64_t n, uint64_t qmask, const std::string &tag, hipStream_t s);

// dev: the states are in device memory (checked there, in row order as on the host)
int validate_problem(const ProblemIn &q, int k, bool dev, hipStream_t st) {
    const std::string tag = "problem " + std::to_string(k) + ": ";
    if (q.n_qubits < 1 || q.n_qubits > 64) return fail(RQX_ERR_ARG, tag + "n_qubits must be in [1, 64]");
    if (q.n_groups < 0 || q.n_terms < 0 || q.n_states < 0) return fail(RQX_ERR_ARG, tag + "negative size");
    if (q.n_groups > 0 && (!q.xmask || !q.grp_off)) return fail(RQX_ERR_ARG, tag + "xmask/grp_off is NULL");
    if (q.n_terms > 0 && (!q.zmask || !q.coef)) return fail(RQX_ERR_ARG, tag + "zmask/coef is NULL");
    if (q.n_states > 0 && !q.states) return fail(RQX_ERR_ARG, tag + "sorted_states is NULL");
    const uint64_t qmask = (q.n_qubits == 64) ? ~0ull : ((1ull << q.n_qubits) - 1ull);
    if (q.n_groups > 0) {
        if (q.grp_off[0] != 0 || q.grp_off[q.n_groups] != q.n_terms)
            return fail(RQX_ERR_ARG, tag + "grp_off must start at 0 and end at n_terms");
        for (int32_t j = 0; j < q.n_groups; ++j) {
            if (q.grp_off[j + 1] < q.grp_off[j]) return fail(RQX_ERR_ARG, tag + "grp_off must be non-decreasing");
            if (q.xmask[j] & ~qmask) return fail(RQX_ERR_ARG, tag + "xmask has bits beyond n_qubits");
            if (j > 0 && q.xmask[j] <= q.xmask[j - 1])
                return fail(RQX_ERR_ARG, tag + "xmask must be strictly ascending (one group per X mask)");
        }
    } else if (q.n_terms != 0) {
        return fail(RQX_ERR_ARG, tag + "terms given without groups");
    }
    for (int64_t t = 0; t < q.n_terms; ++t)
        if (q.zmask[t] & ~qmask) return fail(RQX_ERR_ARG, tag + "zmask has bits beyond n_qubits");
    if (dev) return q.n_states > 0 ? dev_check_states(q.states, q.n_states, qmask, tag, st) : RQX_OK;
    // the first bad state in row order (chunks on host threads for large subspaces)
    const unsigned nt = index_threads(q.n_states);
    std::vector<int> bad_kind(nt, 0);   // per chunk: its first bad state's fault
    parallel_chunks(q.n_states, nt, [&](unsigned w, int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            const int kind = (q.states[i] & ~qmask) ? 1 : (i > 0 && q.states[i] <= q.states[i - 1]) ? 2 : 0;
            if (kind) {
                bad_kind[w] = kind;
                return;
            }
        }
    });
    for (unsigned w = 0; w < nt; ++w) {
        if (bad_kind[w] == 1) return fail(RQX_ERR_ARG, tag + "state has bits beyond n_qubits");
        if (bad_kind[w] == 2) return fail(RQX_ERR_ARG, tag + "states must be sorted ascending and unique");
    }
    return RQX_OK;
}

// Split choice: largest L whose segments fit SEG_CAP, then the cheapest of the
// candidates below it by the sampled cost model (or the caller's L).
// max_segment over chunks: each chunk's first and last segment start and its
// longest segment between them, joined in order
int max_segment_par(const std::vector<int8_t> &hd, int L, unsigned nt) {
    const int64_t n = static_cast<int64_t>(hd.size());
    if (nt <= 1 || n < 2) return max_segment(hd, L);
    std::vector<int64_t> first(nt, -1), last(nt, -1), inner(nt, 0);
    parallel_chunks(n, nt, [&](unsigned w, int64_t b, int64_t e) {
        int64_t f = -1, l = -1, m = 0;
        for (int64_t i = b; i < e; ++i) {
            if (i == 0 || hd[static_cast<size_t>(i)] >= L) {   // row i starts a segment
                if (l >= 0) m = std::max(m, i - l);
                if (f < 0) f = i;
                l = i;
            }
        }
        first[w] = f;
        last[w] = l;
        inner[w] = m;
    });
    int64_t best = 0, prev = -1;
    for (unsigned w = 0; w < nt; ++w) {
        best = std::max(best, inner[w]);
        if (first[w] < 0) continue;
        if (prev >= 0) best = std::max(best, first[w] - prev);
        prev = last[w];
    }
    best = std::max(best, n - prev);
    return static_cast<int>(std::min<int64_t>(best, INT32_MAX));
}

// Section split of the hit cache: the largest b in (L, n_qubits) at which the sorted
// states are an exact product -- every block of equal s >> b holds the same number
// of states with the same low parts s & (2^b - 1) (a Cartesian product of high and
// low strings, e.g. beta x alpha determinants) -- else L.  Groups flipping no bit
// below b then map every row of a block to the same row of another block, so their
// cache levels are row offsets.
int32_t section_split(const uint64_t *st, int64_t n, int32_t L, int32_t n_qubits, int64_t &rows_lo) {
    rows_lo = 0;
    if (n < 2) return L;
    for (int32_t b = n_qubits - 1; b > L; --b) {
        const uint64_t lm = (b >= 64) ? ~0ull : ((1ull << b) - 1ull);
        int64_t m = 1;
        while (m < n && (st[m] >> b) == (st[0] >> b)) ++m;
        if (m == n || n % m != 0) continue;
        bool ok = true;
        for (int64_t blk = m; blk < n && ok; blk += m) {
            if ((st[blk] >> b) == (st[blk - 1] >> b) || (st[blk + m - 1] >> b) != (st[blk] >> b)) {
                ok = false;
                break;
            }
            for (int64_t j = 0; j < m; ++j)
                if ((st[blk + j] & lm) != (st[j] & lm)) {
                    ok = false;
                    break;
                }
        }
        if (ok) {
            rows_lo = m;
            return b;
        }
    }
    return L;
}

// the splits the cost model compares: the nine widest that keep segments within
// SEG_CAP rows (lo_l the widest), then, for a product at a_bits with lists, the
// four below a_bits (the lists need L < a_bits), in descending order
std::vector<int> split_candidates(int lo_l, const AlphaModel &am) {
    std::vector<int> Ls;
    for (int L = lo_l; L >= std::max(0, lo_l - 8); --L) Ls.push_back(L);
    if (!am.listed.empty())
        for (int L = std::min(am.a_bits - 1, lo_l - 9); L >= std::max(0, am.a_bits - 4); --L) Ls.push_back(L);
    return Ls;
}

int choose_split(const ProblemIn &q, int32_t split_bits, bool has_diag, Segmentation &best, int k) {
    const unsigned nt = index_threads(q.n_states);
    std::vector<int8_t> hd(static_cast<size_t>(q.n_states), 64);
    parallel_chunks(q.n_states, nt, [&](unsigned, int64_t b, int64_t e) {
        for (int64_t i = std::max<int64_t>(b, 1); i < e; ++i)
            hd[static_cast<size_t>(i)] = static_cast<int8_t>(highest_diff_bit(q.states[i - 1], q.states[i]));
    });
    const int Lmax_bits = std::min(32, q.n_qubits);
    const std::string tag = "problem " + std::to_string(k) + ": ";
    // the product split of the states, for the alpha lists of the splits below it
    int64_t n_lo = 0;
    const int32_t a_bits = section_split(q.states, q.n_states, 0, q.n_qubits, n_lo);
    const AlphaModel am = alpha_model(q, q.states, n_lo, a_bits);
    if (split_bits >= 0) {
        if (split_bits > Lmax_bits) return fail(RQX_ERR_ARG, tag + "split_bits must be <= min(32, n_qubits)");
        if (max_segment_par(hd, split_bits, nt) > SEG_CAP)
            return fail(RQX_ERR_LIMIT, tag + "split_bits gives a segment larger than " + std::to_string(SEG_CAP) + " rows");
        best.L = split_bits;
        build_segments(best, q.states, q.n_states, q.xmask, q.n_groups);
        best.cost = estimate_cost(best, has_diag, am);
        return RQX_OK;
    }
    int lo_l = 0, hi_l = Lmax_bits;
    while (lo_l < hi_l) {
        const int mid = (lo_l + hi_l + 1) / 2;
        if (max_segment_par(hd, mid, nt) <= SEG_CAP) lo_l = mid; else hi_l = mid - 1;
    }
    // candidate splits, each segmented and costed on its own host thread for large
    // subspaces; the first cheapest (largest L) wins
    const std::vector<int> Ls = split_candidates(lo_l, am);
    const int n_cand = static_cast<int>(Ls.size());
    std::vector<Segmentation> cand(static_cast<size_t>(n_cand));
    auto eval = [&](int c) {
        cand[c].L = Ls[c];
        build_segments(cand[c], q.states, q.n_states, q.xmask, q.n_groups);
        cand[c].cost = estimate_cost(cand[c], has_diag, am);
    };
    if (nt > 1) {
        std::vector<std::thread> pool;
        for (int c = 1; c < n_cand; ++c) pool.emplace_back(eval, c);
        eval(0);
        for (std::thread &t : pool) t.join();
    } else {
        for (int c = 0; c < n_cand; ++c) eval(c);
    }
    int bi = 0;
    for (int c = 1; c < n_cand; ++c)
        if (cand[c].cost < cand[bi].cost) bi = c;
    best = std::move(cand[bi]);
    return RQX_OK;
}

// ---------------------------------------------------------------- device ingest
// The O(N) passes of the index build over states already in device memory
// (straight from the device sort/unique): validation, the segment statistics of
// every split, the segment starts of the candidate splits, the section split,
// the low keys and the cuckoo tables.  Only per-segment arrays (O(M) words) come
// back to the host, where the cost model and the tile lists run as in the host
// build, so both builds give the same plan.

constexpr int ING_WG = 256;      // threads per workgroup of the ingest kernels
constexpr int ING_ROWS = 16;     // rows per thread of k_level_stats
constexpr int ING_LEVELS = 33;   // splits 0..32

// the first bad state: the minimum over rows i of 2 i (bits beyond n_qubits) or
// 2 i + 1 (not above its predecessor), the host check's first fault in row order
__global__ void k_check_states(const uint64_t *__restrict__ st, int64_t n, uint64_t qmask,
                               unsigned long long *first_bad) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t s = st[i];
    unsigned long long code = ~0ull;
    if (s & ~qmask)
        code = 2ull * static_cast<unsigned long long>(i);
    else if (i > 0 && s <= st[i - 1])
        code = 2ull * static_cast<unsigned long long>(i) + 1ull;
    if (code != ~0ull) atomicMin(first_bad, code);
}

// the segment starts of one split over a run of rows: first and last start
// (-1: none), the longest segment between two starts of the run, the starts
struct SegRun {
    int64_t f, l, m, c;
};

__device__ inline SegRun seg_join(const SegRun &a, const SegRun &b) {   // a's rows precede b's
    if (a.c == 0) return b;
    if (b.c == 0) return a;
    const int64_t m = a.m > b.m ? a.m : b.m, gap = b.f - a.l;
    return SegRun{a.f, b.l, m > gap ? m : gap, a.c + b.c};
}

// in-order join over the ING_WG threads of a workgroup (the result in thread 0)
__device__ SegRun seg_reduce(SegRun r, SegRun *red) {
    for (int off = 1; off < 64; off <<= 1) {
        SegRun o;
        o.f = __shfl_down(r.f, off);
        o.l = __shfl_down(r.l, off);
        o.m = __shfl_down(r.m, off);
        o.c = __shfl_down(r.c, off);
        if ((threadIdx.x & 63) + off < 64) r = seg_join(r, o);
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < ING_WG / 64; ++w) r = seg_join(r, red[w]);
    __syncthreads();   // red is reused by the next call
    return r;
}

// per workgroup of ING_WG * ING_ROWS rows: its run of every split L < ING_LEVELS
// (part[L * nwg + wg]; row i starts a segment of split L when i == 0 or its
// highest bit differing from row i - 1 is >= L), and per highest differing bit
// h the first row with that bit (first_pos[h], atomic minimum)
__global__ __launch_bounds__(ING_WG) void k_level_stats(const uint64_t *__restrict__ st, int64_t n, int64_t nwg,
                                                        SegRun *__restrict__ part, unsigned long long *first_pos) {
    __shared__ SegRun red[ING_WG / 64];
    __shared__ unsigned long long fpos[64];
    if (threadIdx.x < 64) fpos[threadIdx.x] = ~0ull;
    const int64_t i0 = (static_cast<int64_t>(blockIdx.x) * ING_WG + threadIdx.x) * ING_ROWS;
    int hd[ING_ROWS];   // 64: row 0, -1: past the end
    uint64_t prev = (i0 > 0 && i0 < n) ? st[i0 - 1] : 0ull;
#pragma unroll
    for (int r = 0; r < ING_ROWS; ++r) {
        const int64_t i = i0 + r;
        hd[r] = -1;
        if (i < n) {
            const uint64_t s = st[i];
            hd[r] = (i == 0) ? 64 : 63 - __clzll(static_cast<long long>(prev ^ s));
            prev = s;
        }
    }
    __syncthreads();   // fpos initialised
    uint64_t seen = 0;
#pragma unroll
    for (int r = 0; r < ING_ROWS; ++r)
        if (hd[r] >= 0 && hd[r] < 64 && !((seen >> hd[r]) & 1ull)) {
            seen |= 1ull << hd[r];
            atomicMin(&fpos[hd[r]], static_cast<unsigned long long>(i0 + r));
        }
    for (int L = 0; L < ING_LEVELS; ++L) {
        SegRun run{-1, -1, 0, 0};
#pragma unroll
        for (int r = 0; r < ING_ROWS; ++r)
            if (hd[r] >= L) {
                const int64_t i = i0 + r;
                if (run.c)
                    run.m = (i - run.l > run.m) ? i - run.l : run.m;
                else
                    run.f = i;
                run.l = i;
                ++run.c;
            }
        run = seg_reduce(run, red);   // synchronises the workgroup (fpos complete after the first)
        if (threadIdx.x == 0) part[static_cast<int64_t>(L) * nwg + blockIdx.x] = run;
    }
    if (threadIdx.x < 64 && fpos[threadIdx.x] != ~0ull) atomicMin(&first_pos[threadIdx.x], fpos[threadIdx.x]);
}

// per split L (one workgroup each): the longest segment and the number of
// segments, from the workgroups' runs joined in row order (out[2 L], out[2 L + 1])
__global__ __launch_bounds__(ING_WG) void k_level_join(const SegRun *__restrict__ part, int64_t nwg, int64_t n,
                                                       int64_t *__restrict__ out) {
    __shared__ SegRun red[ING_WG / 64];
    const int L = blockIdx.x;
    const int64_t per = (nwg + ING_WG - 1) / ING_WG;
    const int64_t w0 = static_cast<int64_t>(threadIdx.x) * per, w1 = (w0 + per < nwg) ? w0 + per : nwg;
    SegRun run{-1, -1, 0, 0};
    for (int64_t w = w0; w < w1; ++w) run = seg_join(run, part[static_cast<int64_t>(L) * nwg + w]);
    run = seg_reduce(run, red);
    if (threadIdx.x == 0) {
        out[2 * L] = run.c ? (run.m > n - run.l ? run.m : n - run.l) : 0;
        out[2 * L + 1] = run.c;
    }
}

// the segment starts of split L, for the device select
struct SegStartAt {
    const uint64_t *st;
    int L;
    __host__ __device__ bool operator()(const uint32_t &i) const {
        return i == 0 || (st[i] >> L) != (st[i - 1] >> L);
    }
};

// seg_hi of the first min(M, *num) selected starts (num: the select's count)
__global__ void k_seg_hi(const uint64_t *__restrict__ st, const uint32_t *__restrict__ seg_off, int64_t M,
                         const int64_t *__restrict__ num, int L, uint64_t *__restrict__ seg_hi) {
    const int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (m < M && m < *num) seg_hi[m] = st[seg_off[m]] >> L;
}

struct SplitCands {   // candidate section splits: bit b[j], first-block length m[j]
    int32_t n;
    int32_t b[64];
    uint32_t m[64];
};

// bit j of *fail: the states are no product at candidate j -- blocks of m[j]
// rows with one value of s >> b[j] each, changing at every block start, whose
// low parts s & (2^b[j] - 1) equal the first block's row for row
__global__ void k_product_check(const uint64_t *__restrict__ st, int64_t n, SplitCands c,
                                unsigned long long *fail) {
    const int64_t i64 = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i64 >= n) return;
    const uint32_t i = static_cast<uint32_t>(i64);
    const uint64_t s = st[i], prev = i > 0 ? st[i - 1] : 0ull;
    unsigned long long bad = 0;
    for (int j = 0; j < c.n; ++j) {
        const int b = c.b[j];
        const uint32_t m = c.m[j], r = i % m;
        const bool same_hi = i > 0 && (s >> b) == (prev >> b);
        const uint64_t lm = (1ull << b) - 1ull;
        if ((r == 0) == same_hi || (i >= m && (s & lm) != (st[r] & lm))) bad |= 1ull << j;
    }
    if (bad) atomicOr(fail, bad);
}

__global__ void k_low_keys(const uint64_t *__restrict__ st, int64_t n, uint32_t lomask, uint32_t *__restrict__ lo) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) lo[i] = static_cast<uint32_t>(st[i]) & lomask;
}

// one thread per segment m with lg[m] >= 0: its cuckoo table of 2 << lg[m]
// entries at tab + off[m]; ok[m] = 1 when the insertions settled
__global__ void k_cuckoo(const uint32_t *__restrict__ lo, const uint32_t *__restrict__ seg_off, int64_t M,
                         const int8_t *__restrict__ lg, const uint64_t *__restrict__ off, uint32_t *tab,
                         uint8_t *__restrict__ ok) {
    const int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (m >= M || lg[m] < 0) return;
    const uint32_t r0 = seg_off[m];
    ok[m] = cuckoo_build(lo + r0, seg_off[m + 1] - r0, static_cast<uint32_t>(lg[m]), tab + off[m]) ? 1 : 0;
}

// packed rows (big endian, pad bit first: sqd.py:255) -> uint64 keys and a
// real-row flag (pad bit 0)
__global__ void k_unpack_rows(const uint8_t *__restrict__ rows, int64_t n, int row_bytes, int n_qubits,
                              uint64_t *__restrict__ keys, uint8_t *__restrict__ real) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint8_t *r = rows + i * row_bytes;
    uint64_t v = 0;   // the first (up to) 8 bytes, left aligned
    for (int k = 0; k < row_bytes && k < 8; ++k) v |= static_cast<uint64_t>(r[k]) << (56 - 8 * k);
    uint64_t key;
    if (row_bytes == 9)   // 64 qubits: the last key bit is the top bit of byte 8
        key = (v << 1) | (static_cast<uint64_t>(r[8]) >> 7);
    else
        key = (v << 1) >> (64 - n_qubits);
    keys[i] = key;
    real[i] = (r[0] >> 7) == 0 ? 1 : 0;
}

inline unsigned grid_for(int64_t n, int wg = ING_WG) { return static_cast<unsigned>((n + wg - 1) / wg); }

// device scratch of the plan build, released on scope exit (after a stream sync
// by its user)
struct DevBuf {
    void *p = nullptr;
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, std::max<size_t>(bytes, 1)); }
    template <typename T> T *at(size_t byte_off = 0) const {
        return reinterpret_cast<T *>(static_cast<char *>(p) + byte_off);
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

#define ING_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(RQX_ERR_HIP, std::string("device ingest: ") + hipGetErrorString(e_)); \
    } while (0)

// the statistics of one problem's device states
struct DevIngest {
    const uint64_t *st = nullptr;
    int64_t n = 0;
    hipStream_t s = nullptr;
    int64_t max_seg[ING_LEVELS] = {};   // longest segment of split L
    int64_t n_seg[ING_LEVELS] = {};     // segments of split L
    int64_t first_pos[64] = {};         // first row whose highest bit differing from its predecessor is h (n: none)
};

// rows beyond n_qubits or out of order, as the host check reports them
int dev_check_states(const uint64_t *st, int64_t n, uint64_t qmask, const std::string &tag, hipStream_t s) {
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, st) != hipSuccess || attr.type != hipMemoryTypeDevice) {
        (void)hipGetLastError();
        return fail(RQX_ERR_ARG, tag + "sorted_states must be device memory");
    }
    DevBuf b;
    ING_TRY(b.alloc(8));
    unsigned long long *bad = b.at<unsigned long long>();
    ING_TRY(fill_async(bad, int64_t(0), int64_t(1), ~0ull, s));
    hipLaunchKernelGGL(k_check_states, dim3(grid_for(n)), dim3(ING_WG), 0, s, st, n, qmask, bad);
    ING_TRY(hipGetLastError());
    unsigned long long h = 0;
    ING_TRY(hipMemcpyAsync(&h, bad, 8, hipMemcpyDeviceToHost, s));
    ING_TRY(hipStreamSynchronize(s));
    if (h == ~0ull) return RQX_OK;
    return fail(RQX_ERR_ARG, tag + ((h & 1) ? "states must be sorted ascending and unique"
                                            : "state has bits beyond n_qubits"));
}

int dev_level_stats(DevIngest &dv) {
    const int64_t nwg = (dv.n + ING_WG * ING_ROWS - 1) / (ING_WG * ING_ROWS);
    const size_t o_part = 0, o_out = align_up(static_cast<size_t>(nwg) * ING_LEVELS * sizeof(SegRun)),
                 o_first = o_out + align_up(2 * ING_LEVELS * 8);
    DevBuf b;
    ING_TRY(b.alloc(o_first + 64 * 8));
    ING_TRY(fill_async(b.at<unsigned long long>(o_first), int64_t(0), int64_t(64), ~0ull, dv.s));
    hipLaunchKernelGGL(k_level_stats, dim3(static_cast<unsigned>(nwg)), dim3(ING_WG), 0, dv.s, dv.st, dv.n, nwg,
                       b.at<SegRun>(o_part), b.at<unsigned long long>(o_first));
    hipLaunchKernelGGL(k_level_join, dim3(ING_LEVELS), dim3(ING_WG), 0, dv.s, b.at<SegRun>(o_part), nwg, dv.n,
                       b.at<int64_t>(o_out));
    ING_TRY(hipGetLastError());
    int64_t out[2 * ING_LEVELS];
    unsigned long long first[64];
    ING_TRY(hipMemcpyAsync(out, b.at<int64_t>(o_out), sizeof(out), hipMemcpyDeviceToHost, dv.s));
    ING_TRY(hipMemcpyAsync(first, b.at<unsigned long long>(o_first), sizeof(first), hipMemcpyDeviceToHost, dv.s));
    ING_TRY(hipStreamSynchronize(dv.s));
    for (int L = 0; L < ING_LEVELS; ++L) {
        dv.max_seg[L] = out[2 * L];
        dv.n_seg[L] = out[2 * L + 1];
    }
    for (int h = 0; h < 64; ++h) dv.first_pos[h] = first[h] == ~0ull ? dv.n : static_cast<int64_t>(first[h]);
    return RQX_OK;
}

// split L's segments from the device states (their starts by a device select),
// then (cost) the classes and the cost model on the host, as build_segments does
int dev_segmentation(const DevIngest &dv, int L, const ProblemIn &q, bool has_diag, Segmentation &sg,
                     const AlphaModel &am, bool cost = true) {
    const int64_t M = dv.n_seg[L];
    if (M < 1 || M > dv.n) return fail(RQX_ERR_HIP, "device ingest: bad segment count");
    const int ni = static_cast<int>(dv.n);
    size_t t_sel = 0;
    hipcub::CountingInputIterator<uint32_t> rows(0);
    ING_TRY(hipcub::DeviceSelect::If(nullptr, t_sel, rows, static_cast<uint32_t *>(nullptr),
                                     static_cast<int64_t *>(nullptr), ni, SegStartAt{dv.st, L}, dv.s));
    // the starts buffer holds n entries, whatever the select finds
    const size_t o_off = align_up(t_sel), o_hi = o_off + align_up(static_cast<size_t>(dv.n) * 4),
                 o_num = o_hi + align_up(static_cast<size_t>(M) * 8);
    DevBuf b;
    ING_TRY(b.alloc(o_num + 8));
    ING_TRY(hipcub::DeviceSelect::If(b.p, t_sel, rows, b.at<uint32_t>(o_off), b.at<int64_t>(o_num), ni,
                                     SegStartAt{dv.st, L}, dv.s));
    hipLaunchKernelGGL(k_seg_hi, dim3(grid_for(M)), dim3(ING_WG), 0, dv.s, dv.st, b.at<uint32_t>(o_off), M,
                       b.at<int64_t>(o_num), L, b.at<uint64_t>(o_hi));
    ING_TRY(hipGetLastError());
    sg.L = L;
    sg.seg_off.assign(static_cast<size_t>(M) + 1, static_cast<uint32_t>(dv.n));
    sg.seg_hi.resize(static_cast<size_t>(M));
    int64_t num = 0;
    ING_TRY(hipMemcpyAsync(sg.seg_off.data(), b.at<uint32_t>(o_off), static_cast<size_t>(M) * 4,
                           hipMemcpyDeviceToHost, dv.s));
    ING_TRY(hipMemcpyAsync(sg.seg_hi.data(), b.at<uint64_t>(o_hi), static_cast<size_t>(M) * 8,
                           hipMemcpyDeviceToHost, dv.s));
    ING_TRY(hipMemcpyAsync(&num, b.at<int64_t>(o_num), 8, hipMemcpyDeviceToHost, dv.s));
    ING_TRY(hipStreamSynchronize(dv.s));
    if (num != M) return fail(RQX_ERR_HIP, "device ingest: segment count mismatch");
    if (cost) {
        segment_classes(sg, q.xmask, q.n_groups);
        sg.cost = estimate_cost(sg, has_diag, am);
    }
    return RQX_OK;
}

int dev_section_split(const DevIngest &dv, int32_t L, int32_t n_qubits, int32_t &a_bits, int64_t &rows_lo);

// choose_split over device states: the same search and candidates, with the
// longest segments of every split from one pass
int dev_choose_split(const DevIngest &dv, const ProblemIn &q, int32_t split_bits, bool has_diag, Segmentation &best,
                     int k) {
    const int Lmax_bits = std::min(32, q.n_qubits);
    const std::string tag = "problem " + std::to_string(k) + ": ";
    // the product split of the states (and its first block, to the host), for
    // the alpha lists of the splits below it
    AlphaModel am;
    {
        int32_t a_bits = 0;
        int64_t n_lo = 0;
        if (int rc = dev_section_split(dv, 0, q.n_qubits, a_bits, n_lo)) return rc;
        if (n_lo > 0) {
            std::vector<uint64_t> first(static_cast<size_t>(n_lo));
            ING_TRY(hipMemcpyAsync(first.data(), dv.st, static_cast<size_t>(n_lo) * 8, hipMemcpyDeviceToHost, dv.s));
            ING_TRY(hipStreamSynchronize(dv.s));
            am = alpha_model(q, first.data(), n_lo, a_bits);
        }
    }
    if (split_bits >= 0) {
        if (split_bits > Lmax_bits) return fail(RQX_ERR_ARG, tag + "split_bits must be <= min(32, n_qubits)");
        if (dv.max_seg[split_bits] > SEG_CAP)
            return fail(RQX_ERR_LIMIT, tag + "split_bits gives a segment larger than " + std::to_string(SEG_CAP) + " rows");
        return dev_segmentation(dv, split_bits, q, has_diag, best, am);
    }
    int lo_l = 0, hi_l = Lmax_bits;
    while (lo_l < hi_l) {
        const int mid = (lo_l + hi_l + 1) / 2;
        if (dv.max_seg[mid] <= SEG_CAP) lo_l = mid; else hi_l = mid - 1;
    }
    // every candidate's segments from the device, then the classes and the cost
    // model of each on its own host thread (as choose_split); the first cheapest
    // (largest L) wins
    const std::vector<int> Ls = split_candidates(lo_l, am);
    const int n_cand = static_cast<int>(Ls.size());
    std::vector<Segmentation> cand(static_cast<size_t>(n_cand));
    for (int c = 0; c < n_cand; ++c)
        if (int rc = dev_segmentation(dv, Ls[c], q, has_diag, cand[c], am, false)) return rc;
    auto eval = [&](int c) {
        segment_classes(cand[c], q.xmask, q.n_groups);
        cand[c].cost = estimate_cost(cand[c], has_diag, am);
    };
    if (index_threads(q.n_states) > 1) {
        std::vector<std::thread> pool;
        for (int c = 1; c < n_cand; ++c) pool.emplace_back(eval, c);
        eval(0);
        for (std::thread &t : pool) t.join();
    } else {
        for (int c = 0; c < n_cand; ++c) eval(c);
    }
    int bi = 0;
    for (int c = 1; c < n_cand; ++c)
        if (cand[c].cost < cand[bi].cost) bi = c;
    best = std::move(cand[bi]);
    return RQX_OK;
}

// section_split over device states: the candidate bits whose first block
// divides the rows, checked in one pass (the largest that holds wins)
int dev_section_split(const DevIngest &dv, int32_t L, int32_t n_qubits, int32_t &a_bits, int64_t &rows_lo) {
    a_bits = L;
    rows_lo = 0;
    if (dv.n < 2) return RQX_OK;
    SplitCands c{};
    for (int32_t b = n_qubits - 1; b > L; --b) {
        int64_t m = dv.n;
        for (int h = b; h < 64; ++h) m = std::min(m, dv.first_pos[h]);
        if (m == dv.n || dv.n % m != 0) continue;
        c.b[c.n] = b;
        c.m[c.n] = static_cast<uint32_t>(m);
        ++c.n;
    }
    if (c.n == 0) return RQX_OK;
    DevBuf b;
    ING_TRY(b.alloc(8));
    ING_TRY(fill_async(b.at<unsigned long long>(), int64_t(0), int64_t(1), 0ull, dv.s));
    hipLaunchKernelGGL(k_product_check, dim3(grid_for(dv.n)), dim3(ING_WG), 0, dv.s, dv.st, dv.n, c,
                       b.at<unsigned long long>());
    ING_TRY(hipGetLastError());
    unsigned long long failed = 0;
    ING_TRY(hipMemcpyAsync(&failed, b.at<unsigned long long>(), 8, hipMemcpyDeviceToHost, dv.s));
    ING_TRY(hipStreamSynchronize(dv.s));
    for (int j = 0; j < c.n; ++j)
        if (!((failed >> j) & 1ull)) {
            a_bits = c.b[j];
            rows_lo = c.m[j];
            break;
        }
    return RQX_OK;
}

// build_hash_tables over the plan's device low keys: every segment tries its
// table sizes in the host order, in rounds (one launch each); a round-0 success
// everywhere leaves the tables in place, else all are rebuilt at their final
// sizes.  hashed = false when some segment does not settle within max_words.
int dev_hash_tables(rqx_plan *p, const std::vector<uint32_t> &seg_off, size_t max_words, hipStream_t s,
                    bool &hashed) {
    hashed = false;
    const size_t M = seg_off.size() - 1;
    if (M == 0) {   // no segment: empty tables, as on the host
        hashed = true;
        return RQX_OK;
    }
    std::vector<int8_t> lg0(M), fin(M, -1);
    for (size_t m = 0; m < M; ++m) {
        const uint32_t n = seg_off[m + 1] - seg_off[m];
        uint32_t lg = 1;
        while ((2u << lg) * 7u < n * 10u) ++lg;   // load <= 0.7
        lg0[m] = static_cast<int8_t>(lg);
    }
    // the hash block: h_off, h_log (M words each), then the tables
    const size_t o_log = align_up(M * 4), o_tab = o_log + align_up(M * 4);
    auto layout = [&](const std::vector<int8_t> &lg, std::vector<uint64_t> &off) -> uint64_t {
        uint64_t total = 0;
        off.assign(M, 0);
        for (size_t m = 0; m < M; ++m) {
            off[m] = total;
            if (lg[m] >= 0) total += uint64_t(2) << lg[m];
        }
        return total;
    };
    DevBuf meta;   // per segment: lg (1 byte), off (8 bytes), ok (1 byte)
    const size_t o_moff = align_up(M), o_mok = o_moff + align_up(M * 8);
    ING_TRY(meta.alloc(o_mok + M));
    auto run = [&](const std::vector<int8_t> &lg, uint32_t *tab, std::vector<uint8_t> &ok) -> hipError_t {
        std::vector<uint64_t> off;
        layout(lg, off);
        ok.assign(M, 0);
        hipError_t e = hipMemcpyAsync(meta.at<int8_t>(), lg.data(), M, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(meta.at<uint64_t>(o_moff), off.data(), M * 8, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_cuckoo, dim3(grid_for(static_cast<int64_t>(M), 64)), dim3(64), 0, s, p->lo,
                               p->seg_off, static_cast<int64_t>(M), meta.at<int8_t>(), meta.at<uint64_t>(o_moff), tab,
                               meta.at<uint8_t>(o_mok));
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(ok.data(), meta.at<uint8_t>(o_mok), M, hipMemcpyDeviceToHost, s);
        const hipError_t es = hipStreamSynchronize(s);   // lg and off are read before they go
        return e == hipSuccess ? es : e;
    };
    std::vector<uint64_t> off0;
    const uint64_t total0 = layout(lg0, off0);
    for (size_t m = 0; m < M; ++m)
        if ((size_t(2) << lg0[m]) > max_words) return RQX_OK;
    if (total0 > UINT32_MAX) return RQX_OK;
    void *blk = nullptr;
    ING_TRY(hipMalloc(&blk, o_tab + std::max<uint64_t>(total0, 1) * 4));
    auto drop = [&]() {
        (void)hipStreamSynchronize(s);
        (void)hipFree(blk);
        blk = nullptr;
    };
    std::vector<uint8_t> ok;
    hipError_t e = run(lg0, reinterpret_cast<uint32_t *>(static_cast<char *>(blk) + o_tab), ok);
    if (e != hipSuccess) {
        drop();
        return fail(RQX_ERR_HIP, std::string("device ingest: ") + hipGetErrorString(e));
    }
    bool all = true;
    for (size_t m = 0; m < M; ++m) {
        if (ok[m]) fin[m] = lg0[m];
        all = all && ok[m];
    }
    if (!all) {
        // further tries of the failed segments (host: up to 3 sizes, lg0 .. lg0 + 2)
        for (int tries = 1; tries < 3; ++tries) {
            std::vector<int8_t> lg(M, -1);
            bool pending = false;
            for (size_t m = 0; m < M; ++m)
                if (fin[m] < 0 && (size_t(2) << (lg0[m] + tries)) <= max_words) {
                    lg[m] = static_cast<int8_t>(lg0[m] + tries);
                    pending = true;
                }
            if (!pending) break;
            std::vector<uint64_t> off;
            const uint64_t total = layout(lg, off);
            DevBuf scratch;
            e = scratch.alloc(total * 4);
            if (e == hipSuccess) e = run(lg, scratch.at<uint32_t>(), ok);
            if (e != hipSuccess) {
                drop();
                return fail(RQX_ERR_HIP, std::string("device ingest: ") + hipGetErrorString(e));
            }
            for (size_t m = 0; m < M; ++m)
                if (lg[m] >= 0 && ok[m]) fin[m] = lg[m];
        }
        for (size_t m = 0; m < M; ++m)
            if (fin[m] < 0) {
                drop();
                return RQX_OK;   // some table does not settle: binary search
            }
        // every table at its final size, rebuilt in place (the same tables as the tries)
        std::vector<uint64_t> off;
        const uint64_t total = layout(fin, off);
        drop();
        if (total > UINT32_MAX) return RQX_OK;
        ING_TRY(hipMalloc(&blk, o_tab + std::max<uint64_t>(total, 1) * 4));
        e = run(fin, reinterpret_cast<uint32_t *>(static_cast<char *>(blk) + o_tab), ok);
        for (size_t m = 0; e == hipSuccess && m < M; ++m)
            if (!ok[m]) e = hipErrorUnknown;
        if (e != hipSuccess) {
            drop();
            return fail(RQX_ERR_HIP, std::string("device ingest (hash rebuild): ") + hipGetErrorString(e));
        }
    }
    std::vector<uint64_t> off;
    layout(fin, off);
    std::vector<uint32_t> h_off(M), h_log(M);
    for (size_t m = 0; m < M; ++m) {
        h_off[m] = static_cast<uint32_t>(off[m]);
        h_log[m] = static_cast<uint32_t>(fin[m]);
        p->hash_words = std::max(p->hash_words, 2 << fin[m]);
    }
    char *c = static_cast<char *>(blk);
    e = hipMemcpyAsync(c, h_off.data(), M * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c + o_log, h_log.data(), M * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        drop();
        p->hash_words = 0;
        return fail(RQX_ERR_HIP, std::string("device ingest: ") + hipGetErrorString(e));
    }
    p->hash_block = blk;
    p->h_off = reinterpret_cast<uint32_t *>(c);
    p->h_log = reinterpret_cast<uint32_t *>(c + o_log);
    p->h_tab = reinterpret_cast<uint32_t *>(c + o_tab);
    hashed = true;
    return RQX_OK;
}

// dev: every problem's states are in device memory (the device ingest passes
// replace the host passes over them; the plan is the same)
int build_plan(rqx_plan **out, const std::vector<ProblemIn> &in, int64_t n_total, int32_t split_bits,
               int persistent, hipStream_t st, bool dev = false) {
    *out = nullptr;
    const int K = static_cast<int>(in.size());
    int64_t N = 0, T = 0, J = 0;
    for (int k = 0; k < K; ++k) {
        if (int rc = validate_problem(in[k], k, dev, st)) return rc;
        N += in[k].n_states;
        T += in[k].n_terms;
        J += in[k].n_groups;
    }
    if (n_total < 0) n_total = N;
    if (n_total < N) return fail(RQX_ERR_ARG, "n_total < n_states");
    if (n_total > 0x7fffffffLL) return fail(RQX_ERR_LIMIT, "n_total must be < 2^31");
    if (T > 0x7fffffffLL || J > 0x7fffffffLL) return fail(RQX_ERR_LIMIT, "terms and groups must be < 2^31");

    std::unique_ptr<rqx_plan> p(new rqx_plan());
    HIP_TRY(hipGetDevice(&p->device));
    HIP_TRY(hipDeviceGetAttribute(&p->n_cu, hipDeviceAttributeMultiprocessorCount, p->device));
    p->persistent = persistent;
    p->n_states = N;
    p->n_total = n_total;
    p->n_terms = T;
    p->n_groups = static_cast<int32_t>(J);

    // host arrays of the whole plan
    std::vector<uint32_t> cls_lane_all;   // per class: every (off-diagonal) group lane-evaluable
    std::vector<uint32_t> lo(dev ? 0 : static_cast<size_t>(N)), seg_off_all, cls_goff_all, g_xlo, g_toff, tile_row0, tile_seg,
        tile_prob;
    std::vector<uint64_t> seg_hi_all, cls_hi_all, g_x, zall;   // (keys are uploaded from the caller's arrays)
    std::vector<double> c64;
    std::vector<float2> c32;
    std::vector<ProbDesc> desc(static_cast<size_t>(std::max(K, 1)));
    std::vector<int64_t> prob_rows(static_cast<size_t>(K) + 1, 0);
    std::vector<double> tile_cost;
    std::vector<int32_t> g_tab;    // per group: first sign-table entry, or -1
    std::vector<uint64_t> g_zor;   // per group: union of its terms' Z masks
    std::vector<GroupDesc> gdesc;
    std::vector<uint32_t> al_lg, al_pg, al_pg_off;   // alpha lists: listed groups; per class, its other groups
    uint32_t al_alpha = 0;                            // alpha indices of all problems (each from a multiple of 64)
    std::vector<SplitDesc> al_split;                  // per group: split values
    std::vector<uint32_t> al_rg, al_roff;             // G-table rows: group, first element
    std::vector<uint64_t> al_rz;                      //   and alpha pattern
    int64_t al_gelems = 0;                            // G-table elements
    std::vector<uint32_t> bl_lg, bl_rg, bl_roff;      // beta lists: listed groups; G-table rows: group, first element
    std::vector<uint64_t> bl_rz;                      //   and beta pattern
    uint32_t bl_nboff = 0;                            //   and offsets
    std::vector<int32_t> al_pmap;                     // partner maps of every problem
    std::vector<uint64_t> al_apats, al_bpats;         //   their alpha and beta patterns
    std::vector<uint64_t> al_apz;                     //   the alpha patterns' reference Z parts
    uint32_t al_nboff = 0;                            //   and beta-partner offsets
    int64_t n_tab = 0;
    int64_t row_base = 0, term_base = 0;
    int32_t grp_base = 0;
    for (int k = 0; k < K; ++k) {
        const ProblemIn &q = in[k];
        ProbHost ph;
        ph.n_qubits = q.n_qubits;
        ph.row_base = row_base;
        ph.n_states = q.n_states;
        ph.term_base = term_base;
        ph.n_terms = q.n_terms;
        ph.grp_base = grp_base;
        ph.n_groups = q.n_groups;
        ph.has_diag = (q.n_groups > 0 && q.xmask[0] == 0) ? 1 : 0;
        ph.diag_terms = ph.has_diag ? static_cast<uint32_t>(q.grp_off[1]) : 0u;
        for (int64_t t = 0; t < q.n_terms; ++t)
            if (q.coef[2 * t + 1] != 0.0) ph.is_real = 0;
        for (int32_t j = 0; j < q.n_groups; ++j) {
            const int len = static_cast<int>(q.grp_off[j + 1] - q.grp_off[j]);
            ph.max_group_terms = std::max(ph.max_group_terms, len);
            // groups with a sign table are never evaluated term by term
            if (q.xmask[j] != 0 && !has_sign_table(q, j)) ph.max_offdiag_terms = std::max(ph.max_offdiag_terms, len);
        }
        Segmentation sg;
        if (dev && q.n_states > 0) {
            DevIngest dv;
            dv.st = q.states;
            dv.n = q.n_states;
            dv.s = st;
            if (int rc = dev_level_stats(dv)) return rc;
            if (int rc = dev_choose_split(dv, q, split_bits, ph.has_diag, sg, k)) return rc;
            if (int rc = dev_section_split(dv, sg.L, q.n_qubits, ph.a_bits, ph.a_rows)) return rc;
        } else {
            if (int rc = choose_split(q, split_bits, ph.has_diag, sg, k)) return rc;
            ph.a_bits = section_split(q.states, q.n_states, sg.L, q.n_qubits, ph.a_rows);
        }
        ph.L = sg.L;
        ph.max_seg = sg.max_seg;
        ph.cost = sg.cost;
        ph.seg_base = static_cast<int32_t>(seg_hi_all.size());
        ph.n_segs = static_cast<int32_t>(sg.seg_hi.size());
        ph.cls_base = static_cast<int32_t>(cls_hi_all.size());
        ph.n_classes = static_cast<int32_t>(sg.cls_hi.size());
        const uint32_t lomask = (sg.L >= 32) ? 0xffffffffu : ((1u << sg.L) - 1u);
        if (!dev)
            parallel_chunks(q.n_states, index_threads(q.n_states), [&](unsigned, int64_t b, int64_t e) {
                for (int64_t i = b; i < e; ++i) lo[row_base + i] = static_cast<uint32_t>(q.states[i]) & lomask;
            });
        for (int32_t m = 0; m < ph.n_segs; ++m) {
            seg_hi_all.push_back(sg.seg_hi[m]);
            seg_off_all.push_back(static_cast<uint32_t>(row_base + sg.seg_off[m]));
            for (uint32_t r = sg.seg_off[m]; r < sg.seg_off[m + 1]; r += TILE) {
                tile_row0.push_back(static_cast<uint32_t>(row_base + r));
                tile_seg.push_back(static_cast<uint32_t>(ph.seg_base + m));
                tile_prob.push_back(static_cast<uint32_t>(k));
                ++ph.n_tiles;
            }
        }
        for (int32_t c = 0; c < ph.n_classes; ++c) {
            cls_hi_all.push_back(sg.cls_hi[c]);
            cls_goff_all.push_back(static_cast<uint32_t>(grp_base + sg.cls_goff[c]));
        }
        for (int32_t j = 0; j < q.n_groups; ++j) {
            g_xlo.push_back(static_cast<uint32_t>(q.xmask[j]) & lomask);
            g_x.push_back(q.xmask[j]);
            g_toff.push_back(static_cast<uint32_t>(term_base + q.grp_off[j]));
            uint64_t zor = 0;
            for (int64_t t = q.grp_off[j]; t < q.grp_off[j + 1]; ++t) zor |= q.zmask[t];
            g_zor.push_back(zor);
            if (has_sign_table(q, j) && n_tab < TAB_MAX_TABLES) {
                p->long_groups = 1;         // the full evaluator reads the tables
                g_tab.push_back(static_cast<int32_t>(n_tab));
                n_tab += 1;
            } else {
                g_tab.push_back(-1);
            }
            GroupDesc gd{};
            gd.x = q.xmask[j];
            gd.z0 = q.grp_off[j + 1] > q.grp_off[j] ? q.zmask[q.grp_off[j]] : 0ull;
            gd.k0 = static_cast<uint32_t>(term_base + q.grp_off[j]);
            gd.k1 = static_cast<uint32_t>(term_base + q.grp_off[j + 1]);
            gd.tab = g_tab.back();
            if (gd.tab >= 0) {
                uint32_t nb = 0;
                for (uint64_t xm = gd.x; xm != 0; xm &= xm - 1, ++nb)
                    gd.xpos |= static_cast<uint32_t>(__builtin_ctzll(xm)) << (6 * nb);
                gd.xpos |= nb << 24;
            }
            gdesc.push_back(gd);
        }
        for (int32_t c = 0; c < ph.n_classes; ++c) {
            uint32_t all = 1;
            for (uint32_t j = sg.cls_goff[c]; j < sg.cls_goff[c + 1] && all; ++j)
                if (q.xmask[j] != 0 && g_tab[static_cast<size_t>(grp_base) + j] < 0 &&
                    q.grp_off[j + 1] - q.grp_off[j] >= TP_MIN_TERMS)
                    all = 0;
            cls_lane_all.push_back(all);
        }
        {
            // alpha lists: in a product at the section split (rows b * n_lo + a) the
            // classes below 2^(a_bits - L) flip only alpha bits.  Their groups that a
            // lane evaluates alone and whose Z masks hold only alpha bits are listed
            // per alpha index with their values; those whose terms with beta Z bits
            // have at most SPLIT_MAX alpha parts are listed with split values; the
            // others are searched (al_pg)
            const bool product = ph.a_bits > sg.L && ph.a_bits < 64 && ph.a_rows > 0 && ph.a_rows < q.n_states &&
                                 q.n_states / ph.a_rows <= 0xFFFFFFFFLL;
            const uint64_t amask = product ? (1ull << ph.a_bits) - 1ull : 0ull;
            auto listed = [&](uint32_t j) {
                const size_t gj = static_cast<size_t>(grp_base) + j;
                return q.xmask[j] != 0 && (g_tab[gj] >= 0 || q.grp_off[j + 1] - q.grp_off[j] < TP_MIN_TERMS) &&
                       (g_zor[gj] & ~amask) == 0;
            };
            std::vector<uint64_t> pat;
            auto split = [&](uint32_t j) { return q.xmask[j] != 0 && !listed(j) && split_patterns(q, j, amask, pat); };
            int32_t n_acls = 0;
            const size_t lg0 = al_lg.size(), r0 = al_rg.size();
            al_split.resize(al_split.size() + static_cast<size_t>(q.n_groups), SplitDesc{});
            const uint32_t nhi = product ? static_cast<uint32_t>(q.n_states / ph.a_rows) : 0u;
            // the problem's G rows, numbered from 0 (the split words' row field),
            // at al_gelems + b * rows + r for beta block b
            if (product)
                for (; n_acls < ph.n_classes && (sg.cls_hi[n_acls] >> (ph.a_bits - sg.L)) == 0; ++n_acls)
                    for (uint32_t j = sg.cls_goff[n_acls]; j < sg.cls_goff[n_acls + 1]; ++j) {
                        const uint32_t gj = static_cast<uint32_t>(grp_base) + j;
                        if (listed(j)) {
                            al_lg.push_back(gj);
                        } else if (split(j)) {
                            al_lg.push_back(gj);
                            SplitDesc &sd = al_split[gj];
                            sd.split = 1;
                            sd.n_p = static_cast<uint32_t>(pat.size());
                            sd.grow = static_cast<uint32_t>(al_rg.size() - r0);
                            for (size_t u = 0; u < pat.size(); ++u) {
                                sd.zeta[u] = pat[u];
                                al_roff.push_back(static_cast<uint32_t>(al_gelems + static_cast<int64_t>(al_rg.size() - r0)));
                                al_rg.push_back(gj);
                                al_rz.push_back(pat[u]);
                            }
                        }
                    }
            const int64_t nrow = static_cast<int64_t>(al_rg.size() - r0), gel = al_gelems + nrow * nhi;
            if (al_lg.size() == lg0 || gel > 0xFFFFFFFFLL || nrow >= (int64_t(1) << SW_ROW_BITS)) {
                n_acls = 0;
                al_lg.resize(lg0);
                al_rg.resize(r0);
                al_rz.resize(r0);
                al_roff.resize(r0);
                for (int32_t j = 0; j < q.n_groups; ++j) al_split[static_cast<size_t>(grp_base) + j] = SplitDesc{};
            } else {
                ph.al_nlo = static_cast<uint32_t>(ph.a_rows);
                ph.al_nhi = nhi;
                ph.al_a0 = al_alpha;
                ph.al_lg0 = static_cast<uint32_t>(lg0);
                ph.al_lgn = static_cast<uint32_t>(al_lg.size() - lg0);
                ph.al_r0 = static_cast<uint32_t>(r0);
                ph.al_rn = static_cast<uint32_t>(nrow);
                ph.al_g0 = static_cast<uint32_t>(al_gelems);
                al_alpha += (ph.al_nlo + 63u) & ~63u;
                al_gelems = gel;
            }
            for (int32_t c = 0; c < ph.n_classes; ++c) {
                al_pg_off.push_back(static_cast<uint32_t>(al_pg.size()));
                if (c >= n_acls) continue;
                cls_lane_all[static_cast<size_t>(ph.cls_base) + c] |= 2u;
                for (uint32_t j = sg.cls_goff[c]; j < sg.cls_goff[c + 1]; ++j)
                    if (q.xmask[j] != 0 && !listed(j) && !al_split[static_cast<size_t>(grp_base) + j].split)
                        al_pg.push_back(static_cast<uint32_t>(grp_base) + j);
            }
        }
        if (ph.al_nlo > 0 && ph.a_bits <= 32) {
            // the partner map: a class above the alpha classes whose groups a lane
            // evaluates alone (mapped) takes its hits of the groups that flip alpha
            // bits from the row's alpha partners (alpha parts X_a of those groups
            // with alpha ^ x_a an alpha string) and the tile's beta partners (beta
            // parts X_b, likewise), the map giving the group of each (x_b, x_a)
            const uint64_t amask = (1ull << ph.a_bits) - 1ull;
            std::vector<uint64_t> Xa, Xb;
            std::vector<int32_t> mcls;
            for (int32_t c = 0; c < ph.n_classes; ++c) {
                if ((sg.cls_hi[c] >> (ph.a_bits - sg.L)) == 0) continue;                       // an alpha class
                // (a group of high bits only, the class's first, is taken on its own)
                bool lane = true, any = false;
                for (uint32_t j = sg.cls_goff[c]; j < sg.cls_goff[c + 1] && lane; ++j)
                    if ((q.xmask[j] & amask) != 0) {
                        any = true;
                        lane = g_tab[static_cast<size_t>(grp_base) + j] >= 0 ||
                               q.grp_off[j + 1] - q.grp_off[j] < TP_MIN_TERMS;
                    }
                if (!any || !lane) continue;
                for (uint32_t j = sg.cls_goff[c]; j < sg.cls_goff[c + 1]; ++j)
                    if ((q.xmask[j] & amask) != 0) {
                        Xa.push_back(q.xmask[j] & amask);
                        Xb.push_back(q.xmask[j] & ~amask);
                    }
                mcls.push_back(c);
            }
            for (std::vector<uint64_t> *X : {&Xa, &Xb}) {
                std::sort(X->begin(), X->end());
                X->erase(std::unique(X->begin(), X->end()), X->end());
            }
            const int64_t elems = static_cast<int64_t>(Xa.size()) * static_cast<int64_t>(Xb.size());
            if (!mcls.empty() && static_cast<int64_t>(al_pmap.size()) + elems <= CMAP_MAX_ELEMS) {
                ph.pm_base = static_cast<uint32_t>(al_pmap.size());
                ph.al_pat0 = static_cast<uint32_t>(al_apats.size());
                ph.al_npat = static_cast<uint32_t>(Xa.size());
                ph.al_bpat0 = static_cast<uint32_t>(al_bpats.size());
                ph.al_nbpat = static_cast<uint32_t>(Xb.size());
                ph.bp_b0 = al_nboff;
                al_nboff += ph.al_nhi + 1;
                al_apats.insert(al_apats.end(), Xa.begin(), Xa.end());
                al_bpats.insert(al_bpats.end(), Xb.begin(), Xb.end());
                al_pmap.resize(al_pmap.size() + static_cast<size_t>(elems), -1);
                // per alpha pattern a reference Z part: the alpha part of the first
                // term of its first group with a sign table.  The product kernel
                // stages, per beta partner, the values of the pattern's table groups
                // whose first terms' alpha parts differ from it only at the pattern's
                // X bits, over those bits (2^popc entries; al_hsl = the largest popc
                // below TAB_MAX_BITS)
                std::vector<uint64_t> pz(Xa.size(), 0);
                std::vector<uint8_t> pz_set(Xa.size(), 0);
                ph.al_hsl = 0;
                for (int32_t c : mcls) {
                    cls_lane_all[static_cast<size_t>(ph.cls_base) + c] |= 4u;
                    for (uint32_t j = sg.cls_goff[c]; j < sg.cls_goff[c + 1]; ++j) {
                        const uint64_t xa = q.xmask[j] & amask;
                        if (xa == 0) continue;
                        const size_t ia = static_cast<size_t>(std::lower_bound(Xa.begin(), Xa.end(), xa) - Xa.begin());
                        const size_t ib = static_cast<size_t>(
                            std::lower_bound(Xb.begin(), Xb.end(), q.xmask[j] & ~amask) - Xb.begin());
                        al_pmap[ph.pm_base + ib * Xa.size() + ia] = static_cast<int32_t>(grp_base) + static_cast<int32_t>(j);
                        const uint32_t na = static_cast<uint32_t>(__builtin_popcountll(xa));
                        if (g_tab[static_cast<size_t>(grp_base) + j] >= 0 && na < static_cast<uint32_t>(TAB_MAX_BITS)) {
                            if (!pz_set[ia]) {
                                pz[ia] = q.zmask[q.grp_off[j]] & amask;
                                pz_set[ia] = 1;
                            }
                            ph.al_hsl = std::max(ph.al_hsl, na);
                        }
                    }
                }
                al_apz.insert(al_apz.end(), pz.begin(), pz.end());
            }
        }
        if (ph.al_nlo > 0 && ph.a_bits <= 32) {
            // beta lists: when the alpha classes search nothing and every other class
            // is mapped or holds only a group of high bits, those groups come from
            // per-beta lists (values, or split values with G tables per alpha string)
            // and the tile kernel searches nothing at all
            const uint64_t amask = (1ull << ph.a_bits) - 1ull;
            bool full = al_pg.size() == al_pg_off[static_cast<size_t>(ph.cls_base)];
            std::vector<uint32_t> blg;
            std::vector<std::vector<uint64_t>> bpat;
            std::vector<uint64_t> pat;
            for (int32_t c = 0; c < ph.n_classes && full; ++c) {
                if ((sg.cls_hi[c] >> (ph.a_bits - sg.L)) == 0) continue;                       // an alpha class
                uint32_t j = sg.cls_goff[c];
                if ((q.xmask[j] & amask) == 0) {         // the class's group of high bits only
                    if (!split_patterns(q, static_cast<int32_t>(j), ~amask, pat)) full = false;
                    blg.push_back(static_cast<uint32_t>(grp_base) + j);
                    bpat.push_back(pat);
                    ++j;
                }
                if (j < sg.cls_goff[c + 1] && !(cls_lane_all[static_cast<size_t>(ph.cls_base) + c] & 4u)) full = false;
            }
            int64_t gel = al_gelems;
            for (const std::vector<uint64_t> &P : bpat) gel += static_cast<int64_t>(P.size()) * ph.al_nlo;
            if (full && !blg.empty() && gel <= 0xFFFFFFFFLL) {
                ph.full = 1;
                ph.bl_b0 = bl_nboff;
                bl_nboff += ph.al_nhi + 1;
                ph.bl_lg0 = static_cast<uint32_t>(bl_lg.size());
                ph.bl_lgn = static_cast<uint32_t>(blg.size());
                ph.bl_r0 = static_cast<uint32_t>(bl_rg.size());
                for (size_t u = 0; u < blg.size(); ++u) {
                    bl_lg.push_back(blg[u]);
                    // split when some term holds alpha Z bits, or when the group is too long
                    // for one lane (F is then its whole sum); else its value as listed
                    const bool lane = g_tab[blg[u]] >= 0 ||
                                      q.grp_off[blg[u] - grp_base + 1] - q.grp_off[blg[u] - grp_base] < TP_MIN_TERMS;
                    if (!bpat[u].empty() || !lane) {        // split (F, and G rows when it has patterns)
                        SplitDesc &sd = al_split[blg[u]];
                        sd.split = 1;
                        sd.n_p = static_cast<uint32_t>(bpat[u].size());
                        sd.grow = static_cast<uint32_t>(al_gelems);
                        for (size_t v = 0; v < bpat[u].size(); ++v) {
                            sd.zeta[v] = bpat[u][v];
                            bl_rg.push_back(blg[u]);
                            bl_rz.push_back(bpat[u][v]);
                            bl_roff.push_back(static_cast<uint32_t>(al_gelems));
                            al_gelems += ph.al_nlo;
                        }
                    }
                }
                ph.bl_rn = static_cast<uint32_t>(bl_rg.size() - ph.bl_r0);
            }
        }
        for (int64_t t = 0; t < q.n_terms; ++t) {
            zall.push_back(q.zmask[t]);
            c64.push_back(q.coef[2 * t]);
            c64.push_back(q.coef[2 * t + 1]);
            c32.push_back(make_float2(static_cast<float>(q.coef[2 * t]), static_cast<float>(q.coef[2 * t + 1])));
        }
        for (int32_t t = 0; t < ph.n_tiles; ++t) tile_cost.push_back(ph.n_tiles ? ph.cost / ph.n_tiles : 0.0);
        desc[k] = ProbDesc{static_cast<uint32_t>(ph.seg_base), static_cast<uint32_t>(ph.n_segs),
                           static_cast<uint32_t>(ph.cls_base), static_cast<uint32_t>(ph.n_classes),
                           static_cast<uint32_t>(ph.grp_base), lomask, ph.has_diag, ph.a_bits};
        desc[k].row0 = static_cast<uint32_t>(row_base);
        desc[k].al_nlo = ph.al_nlo;
        desc[k].al_a0 = ph.al_a0;
        desc[k].al_nhi = ph.al_nhi;
        desc[k].pm_base = ph.pm_base;
        desc[k].pm_na = ph.al_npat;
        desc[k].bp_b0 = ph.bp_b0;
        desc[k].bl_b0 = ph.bl_b0;
        desc[k].full = ph.full;
        desc[k].al_g0 = ph.al_g0;
        desc[k].al_gnr = ph.al_rn;
        desc[k].al_pat0 = ph.al_pat0;
        desc[k].al_hsl = ph.al_hsl;
        desc[k].prod = ph.a_bits > ph.L ? 1u : 0u;
        p->max_seg = std::max(p->max_seg, ph.max_seg);
        p->is_real = p->is_real && ph.is_real;
        if (ph.max_offdiag_terms >= TP_MIN_TERMS) p->long_groups = 1;
        row_base += q.n_states;
        term_base += q.n_terms;
        grp_base += q.n_groups;
        prob_rows[k + 1] = row_base;
        p->probs.push_back(ph);
    }
    // segment and class ranges end with their sentinels
    seg_off_all.push_back(static_cast<uint32_t>(N));
    cls_goff_all.push_back(static_cast<uint32_t>(J));
    al_pg_off.push_back(static_cast<uint32_t>(al_pg.size()));
    g_toff.push_back(static_cast<uint32_t>(T));
    p->n_segs = static_cast<int32_t>(seg_hi_all.size());
    p->n_classes = static_cast<int32_t>(cls_hi_all.size());
    p->n_tiles = static_cast<int32_t>(tile_row0.size());
    if (persistent && p->n_tiles > 1) {
        // longest tiles first (LPT order for the work queue)
        std::vector<uint32_t> order(tile_row0.size());
        for (size_t t = 0; t < order.size(); ++t) order[t] = static_cast<uint32_t>(t);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return tile_cost[a] > tile_cost[b]; });
        std::vector<uint32_t> r0(order.size()), sg(order.size()), pr(order.size());
        for (size_t t = 0; t < order.size(); ++t) {
            r0[t] = tile_row0[order[t]];
            sg[t] = tile_seg[order[t]];
            pr[t] = tile_prob[order[t]];
        }
        tile_row0.swap(r0);
        tile_seg.swap(sg);
        tile_prob.swap(pr);
    }

    p->n_tab = static_cast<uint32_t>(n_tab);

    // hashed target lookups when every problem's low keys fit the entries
    std::vector<uint32_t> h_off, h_log, h_tab;
    bool hash_fits = true;
    for (const ProbHost &ph : p->probs) hash_fits = hash_fits && ph.L <= HASH_MAX_KEY_BITS;
    int lds_max = 65536;
    HIP_TRY(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device));
    p->lds_max = lds_max;
    // the matvec's hit queues share the workgroup's LDS with the staged table
    const size_t hash_max_words = static_cast<size_t>(std::max(lds_max - 256 - static_cast<int>(HQ_LDS_MAX), 0)) / 4;
    if (dev) {
        // built after the upload, from the device low keys
    } else if (hash_fits && !seg_off_all.empty() &&
        build_hash_tables(lo, seg_off_all, hash_max_words, h_off, h_log, h_tab)) {
        p->hashed = 1;
        for (uint32_t lg : h_log) p->hash_words = std::max(p->hash_words, 2 << lg);
    } else {
        h_off.clear();
        h_log.clear();
        h_tab.clear();
    }

    // term arrays padded with zero terms (coefficient sums fetch TERM_PAD terms at a time)
    zall.resize(zall.size() + TERM_PAD, 0ull);
    c64.resize(c64.size() + 2 * TERM_PAD, 0.0);
    c32.resize(c32.size() + TERM_PAD, make_float2(0.f, 0.f));

    // ---- one device block, carved
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t o = off;
        off += align_up(std::max<size_t>(bytes, 1));
        return o;
    };
    const size_t n = static_cast<size_t>(N);
    const size_t o_keys = carve(n * 8), o_lo = carve(n * 4), o_seghi = carve(seg_hi_all.size() * 8),
                 o_segoff = carve(seg_off_all.size() * 4), o_clshi = carve(cls_hi_all.size() * 8),
                 o_clsgoff = carve(cls_goff_all.size() * 4), o_clslane = carve(cls_lane_all.size() * 4),
                 o_gxlo = carve(g_xlo.size() * 4),
                 o_gtoff = carve(g_toff.size() * 4), o_gx = carve(g_x.size() * 8),
                 o_z = carve(zall.size() * 8), o_c64 = carve(c64.size() * 8),
                 o_c32 = carve(c32.size() * 8), o_d64 = carve(n * 16),
                 o_d32 = carve(n * 8), o_trow = carve(tile_row0.size() * 4), o_tseg = carve(tile_seg.size() * 4),
                 o_tprob = carve(tile_prob.size() * 4), o_desc = carve(desc.size() * sizeof(ProbDesc)),
                 o_prows = carve(prob_rows.size() * 8), o_queue = carve(sizeof(unsigned int)),
                 o_gtab = carve(g_tab.size() * 4), o_gzor = carve(g_zor.size() * 8),
                 o_gdesc = carve(gdesc.size() * sizeof(GroupDesc)), o_t64re = carve(static_cast<size_t>(n_tab) * TAB_ENTRIES * 8),
                 o_t64im = carve(static_cast<size_t>(n_tab) * TAB_ENTRIES * 8),
                 o_t32re = carve(static_cast<size_t>(n_tab) * TAB_ENTRIES * 4),
                 o_t32im = carve(static_cast<size_t>(n_tab) * TAB_ENTRIES * 4), o_hoff = carve(h_off.size() * 4),
                 o_hlog = carve(h_log.size() * 4), o_htab = carve(h_tab.size() * 4);
    int64_t max_rows = 0;
    for (const ProbHost &ph : p->probs) max_rows = std::max(max_rows, ph.n_states);
    p->chunks = vector_chunks(max_rows);
    const size_t o_part = carve(static_cast<size_t>(K) * p->chunks * 2 * LP_DOT * 8);
    HIP_TRY(hipMalloc(&p->block, off));
    char *b = static_cast<char *>(p->block);
    p->keys = reinterpret_cast<uint64_t *>(b + o_keys);
    p->lo = reinterpret_cast<uint32_t *>(b + o_lo);
    p->seg_hi = reinterpret_cast<uint64_t *>(b + o_seghi);
    p->seg_off = reinterpret_cast<uint32_t *>(b + o_segoff);
    p->cls_hi = reinterpret_cast<uint64_t *>(b + o_clshi);
    p->cls_goff = reinterpret_cast<uint32_t *>(b + o_clsgoff);
    p->cls_lane = reinterpret_cast<uint32_t *>(b + o_clslane);
    p->g_xlo = reinterpret_cast<uint32_t *>(b + o_gxlo);
    p->g_toff = reinterpret_cast<uint32_t *>(b + o_gtoff);
    p->g_x = reinterpret_cast<uint64_t *>(b + o_gx);
    p->z = reinterpret_cast<uint64_t *>(b + o_z);
    p->coef64 = reinterpret_cast<double2 *>(b + o_c64);
    p->coef32 = reinterpret_cast<float2 *>(b + o_c32);
    p->diag64 = reinterpret_cast<double2 *>(b + o_d64);
    p->diag32 = reinterpret_cast<float2 *>(b + o_d32);
    p->tile_row0 = reinterpret_cast<uint32_t *>(b + o_trow);
    p->tile_seg = reinterpret_cast<uint32_t *>(b + o_tseg);
    p->tile_prob = reinterpret_cast<uint32_t *>(b + o_tprob);
    p->prob = reinterpret_cast<ProbDesc *>(b + o_desc);
    p->prob_rows = reinterpret_cast<int64_t *>(b + o_prows);
    p->queue = reinterpret_cast<unsigned int *>(b + o_queue);
    p->dot_part = reinterpret_cast<double *>(b + o_part);
    p->g_tab = reinterpret_cast<int32_t *>(b + o_gtab);
    p->g_zor = reinterpret_cast<uint64_t *>(b + o_gzor);
    p->gdesc = reinterpret_cast<GroupDesc *>(b + o_gdesc);
    p->tab64re = reinterpret_cast<double *>(b + o_t64re);
    p->tab64im = reinterpret_cast<double *>(b + o_t64im);
    p->tab32re = reinterpret_cast<float *>(b + o_t32re);
    p->tab32im = reinterpret_cast<float *>(b + o_t32im);
    p->h_off = reinterpret_cast<uint32_t *>(b + o_hoff);
    p->h_log = reinterpret_cast<uint32_t *>(b + o_hlog);
    p->h_tab = reinterpret_cast<uint32_t *>(b + o_htab);
    rqx_plan *raw = p.get();
    auto release = [&]() {
        (void)hipStreamSynchronize(st);   // no copy or kernel of the plan still in flight
        (void)hipFree(raw->block);
        (void)hipFree(raw->st_block);
        (void)hipFree(raw->hash_block);
        (void)hipFree(raw->al_block);
        (void)hipFree(raw->al_ent_block);
        (void)hipFree(raw->al_pblock);
        (void)hipFree(raw->al_pent_block);
        (void)hipFree(raw->bl_block);
        (void)hipFree(raw->bl_ent_block);
    };

    auto up = [&](void *dst, const void *src, size_t bytes) -> hipError_t {
        if (bytes == 0) return hipSuccess;
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    };
    const hipError_t errs[] = {
        up(p->lo, lo.data(), lo.size() * 4),
        up(p->seg_hi, seg_hi_all.data(), seg_hi_all.size() * 8), up(p->seg_off, seg_off_all.data(), seg_off_all.size() * 4),
        up(p->cls_hi, cls_hi_all.data(), cls_hi_all.size() * 8), up(p->cls_goff, cls_goff_all.data(), cls_goff_all.size() * 4),
        up(p->cls_lane, cls_lane_all.data(), cls_lane_all.size() * 4),
        up(p->g_xlo, g_xlo.data(), g_xlo.size() * 4), up(p->g_toff, g_toff.data(), g_toff.size() * 4),
        up(p->g_x, g_x.data(), g_x.size() * 8), up(p->z, zall.data(), zall.size() * 8),
        up(p->coef64, c64.data(), c64.size() * 8), up(p->coef32, c32.data(), c32.size() * 8),
        up(p->tile_row0, tile_row0.data(), tile_row0.size() * 4), up(p->tile_seg, tile_seg.data(), tile_seg.size() * 4),
        up(p->tile_prob, tile_prob.data(), tile_prob.size() * 4), up(p->prob, desc.data(), desc.size() * sizeof(ProbDesc)),
        up(p->prob_rows, prob_rows.data(), prob_rows.size() * 8), up(p->g_tab, g_tab.data(), g_tab.size() * 4),
        up(p->g_zor, g_zor.data(), g_zor.size() * 8), up(p->gdesc, gdesc.data(), gdesc.size() * sizeof(GroupDesc)),
        up(p->h_off, h_off.data(), h_off.size() * 4), up(p->h_log, h_log.data(), h_log.size() * 4),
        up(p->h_tab, h_tab.data(), h_tab.size() * 4)};
    std::vector<hipError_t> errs_keys;   // the keys straight from the callers' arrays (read before return)
    for (int k = 0; k < K; ++k) {
        const size_t bytes = static_cast<size_t>(in[k].n_states) * 8;
        errs_keys.push_back(!dev ? up(p->keys + prob_rows[k], in[k].states, bytes)
                            : bytes ? hipMemcpyAsync(p->keys + prob_rows[k], in[k].states, bytes,
                                                     hipMemcpyDeviceToDevice, st)
                                    : hipSuccess);
    }
    for (hipError_t ei : errs_keys)
        if (ei != hipSuccess) {
            release();
            return fail(RQX_ERR_HIP, std::string("hipMemcpyAsync(plan): ") + hipGetErrorString(ei));
        }
    for (hipError_t ei : errs)
        if (ei != hipSuccess) {
            release();
            return fail(RQX_ERR_HIP, std::string("hipMemcpyAsync(plan): ") + hipGetErrorString(ei));
        }
    for (size_t k = 0; k < p->probs.size(); ++k) {
        const ProbHost &ph = p->probs[k];
        if (ph.n_states == 0) continue;
        const unsigned blocks = static_cast<unsigned>((ph.n_states + 255) / 256);
        const uint32_t k0 = static_cast<uint32_t>(ph.term_base);
        hipLaunchKernelGGL(k_diag, dim3(blocks), dim3(256), 0, st, p->keys + ph.row_base, ph.n_states, p->z, p->coef64,
                           k0, k0 + ph.diag_terms, p->diag64 + ph.row_base, p->diag32 + ph.row_base);
        if (dev)
            hipLaunchKernelGGL(k_low_keys, dim3(grid_for(ph.n_states)), dim3(ING_WG), 0, st, p->keys + ph.row_base,
                               ph.n_states, desc[k].lomask, p->lo + ph.row_base);
    }
    if (dev && hash_fits && !seg_off_all.empty()) {
        bool hashed = false;
        if (int rc = dev_hash_tables(raw, seg_off_all, hash_max_words, st, hashed)) {
            release();
            return rc;
        }
        p->hashed = hashed ? 1 : 0;
    }
    launch_group_tables(raw, st);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = build_targets(raw, st);
    if (e == hipSuccess)
        e = build_alpha_lists(raw, al_alpha, al_lg, al_pg_off, al_pg, al_split, al_rg, al_rz, al_roff, al_gelems, desc, st);
    if (e == hipSuccess) e = build_partners(raw, al_alpha, al_nboff, al_apats, al_apz, al_bpats, al_pmap, desc, st);
    if (e == hipSuccess) e = build_beta_lists(raw, bl_nboff, bl_lg, bl_rg, bl_rz, bl_roff, desc, st);
    if (e == hipSuccess) {
        raw->all_full = !raw->probs.empty();
        for (const ProbHost &ph : raw->probs) raw->all_full = raw->all_full && ph.full && ph.n_states > 0;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) {
        int occ = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &occ, reinterpret_cast<const void *>(&k_tiles_queue<MODE_MATVEC, double, true>), TILE,
            ((seg_lds_bytes(raw) + 15) & ~size_t(15)) + HQ_LDS_MAX);
        p->occupancy = std::max(occ, 1);
    }
    if (e != hipSuccess) {
        release();
        return fail(RQX_ERR_HIP, std::string("plan build: ") + hipGetErrorString(e));
    }
    *out = p.release();
    return RQX_OK;
}

int check_dtype(const rqx_plan *p, int32_t dtype) {
    if (dtype < RQX_F64 || dtype > RQX_C64) return fail(RQX_ERR_ARG, "unknown dtype");
    const bool cplx = (dtype == RQX_C128 || dtype == RQX_C64);
    if (!cplx && !p->is_real)
        return fail(RQX_ERR_DTYPE, "real vector dtype requested but the Hamiltonian has complex coefficients");
    return RQX_OK;
}

bool is_fp64(int32_t dtype) { return dtype == RQX_F64 || dtype == RQX_C128; }

int cache_slot(int32_t precision) { return precision == 64 ? 1 : precision == 32 ? 0 : -1; }

size_t cache_val_bytes(int slot, int cplx) { return (slot == 1 ? 8u : 4u) * (cplx ? 2u : 1u); }

hipError_t free_cache(HitCache &c) {
    hipError_t e = hipSuccess;
    for (void *b : {c.block, c.col_block, c.vala_block, c.sell_block}) {
        if (!b) continue;
        const hipError_t e2 = hipFree(b);
        if (e == hipSuccess) e = e2;
    }
    c = HitCache{};
    return e;
}

// release the cache of one slot; caches built together are released together
hipError_t release_cache(rqx_plan *p, int slot) {
    for (int s = 0; s < 2; ++s)
        if (s == slot || p->cache[slot].shared) {
            p->cache_slower[s ? RQX_F64 : RQX_F32] = 0;
            p->cache_slower[s ? RQX_C128 : RQX_C64] = 0;
        }
    if (!p->cache[slot].shared) return free_cache(p->cache[slot]);
    const hipError_t e = free_cache(p->cache[1]);
    p->cache[0] = HitCache{};
    return e;
}

// the column pass: columns in the full packed layout (slice s from off[s]) and
// the group of each section-A entry (level-major), in the order the tile kernel
// accumulates the hits
hipError_t fill_columns(const rqx_plan *p, int slot, const HitCache &c, const int64_t *off, int32_t *col_full,
                        int32_t *gid_a, void *valb_lo, hipStream_t st) {
    KParams k = make_params(p, slot == 1);
    k.c_col = col_full;
    k.c_gid = gid_a;
    k.c_off = off;
    k.c_wa = c.wa;
    k.c_wb = c.wb;
    k.c_bpos = c.bpos;
    k.c_bw = c.bw;
    k.c_aoff = c.aoff;
    k.c_boff = c.boff;
    k.c_val = c.valb;
    if (valb_lo) {
        k.c_val_lo = valb_lo;
        k.coef_lo = p->coef32;
    }
    if (slot == 1)
        return c.cplx ? launch_tiles<MODE_FILLCOL, double, true>(p, k, nullptr, nullptr, st)
                      : launch_tiles<MODE_FILLCOL, double, false>(p, k, nullptr, nullptr, st);
    return c.cplx ? launch_tiles<MODE_FILLCOL, float, true>(p, k, nullptr, nullptr, st)
                  : launch_tiles<MODE_FILLCOL, float, false>(p, k, nullptr, nullptr, st);
}

// the value pass: the values of the cache in slot (both precisions in one pass
// when they were built together) from the current coefficients, straight into
// the stored layout (a_only: section A only, after a column pass); it needs no
// memory beyond the cache
hipError_t fill_values(const rqx_plan *p, int slot, hipStream_t st, bool a_only = false) {
    if (p->cache[slot].shared) slot = 1;
    const HitCache &c = p->cache[slot];
    KParams k = make_params(p, slot == 1);
    k.c_wa = c.wa;
    k.c_wb = c.wb;
    k.c_bpos = c.bpos;
    k.c_bw = c.bw;
    k.c_aoff = c.aoff;
    k.c_boff = c.boff;
    k.c_avoff = c.avoff;
    k.c_acode = c.acode;
    k.c_val = c.valb;
    k.c_aval = c.aval;
    k.c_vala = c.vala;
    k.c_a_only = a_only ? 1 : 0;
    if (c.shared) {
        const HitCache &lo = p->cache[0];
        k.c_val_lo = lo.valb;
        k.c_aval_lo = lo.aval;
        k.c_vala_lo = lo.vala;
        k.coef_lo = p->coef32;
    }
    if (slot == 1)
        return c.cplx ? launch_tiles<MODE_FILL, double, true>(p, k, nullptr, nullptr, st)
                      : launch_tiles<MODE_FILL, double, false>(p, k, nullptr, nullptr, st);
    return c.cplx ? launch_tiles<MODE_FILL, float, true>(p, k, nullptr, nullptr, st)
                  : launch_tiles<MODE_FILL, float, false>(p, k, nullptr, nullptr, st);
}

template <typename R, bool CPLXV>
hipError_t spmv_launch(const rqx_plan *p, const HitCache &c, const void *diag, const void *x, void *y, hipStream_t st) {
    using V = typename VT<R, CPLXV>::T;
    const unsigned blocks = static_cast<unsigned>((p->n_states + SPMV_WG - 1) / SPMV_WG);
    const size_t lds = static_cast<size_t>(std::max(1, std::min(c.win_max, stage_cap<V>()))) * sizeof(V);
    if (c.cplx)
        hipLaunchKernelGGL((k_spmv_sell<R, CPLXV, true>), dim3(blocks), dim3(SPMV_WG), lds, st, c.col, c.valb, c.vala,
                           c.aval, c.boff, c.avoff, c.coff, c.aoff, c.adel, c.acode, c.wa, c.bperm, c.bw, c.win, diag,
                           p->n_states, x, y);
    else
        hipLaunchKernelGGL((k_spmv_sell<R, CPLXV, false>), dim3(blocks), dim3(SPMV_WG), lds, st, c.col, c.valb, c.vala,
                           c.aval, c.boff, c.avoff, c.coff, c.aoff, c.adel, c.acode, c.wa, c.bperm, c.bw, c.win, diag,
                           p->n_states, x, y);
    return hipGetLastError();
}

}  // namespace

extern "C" {

int rqx_version(void) { return 200; }

const char *rqx_last_error(void) { return g_err.c_str(); }

int rqx_plan_create(rqx_plan **out, int32_t n_qubits, const uint64_t *xmask, const int64_t *grp_off,
                    int32_t n_groups, const uint64_t *zmask, const double *coef, int64_t n_terms,
                    const uint64_t *sorted_states, int64_t n_states, int64_t n_total, int32_t split_bits,
                    rqx_stream_t stream) {
    g_err.clear();
    if (!out) return fail(RQX_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (n_total < n_states) return fail(RQX_ERR_ARG, "negative size or n_total < n_states");
    std::vector<ProblemIn> in = {{n_qubits, xmask, grp_off, n_groups, zmask, coef, n_terms, sorted_states, n_states}};
    return build_plan(out, in, n_total, split_bits, 0, static_cast<hipStream_t>(stream));
}

int rqx_plan_create_batch(rqx_plan **out, int32_t n_problems, const int32_t *n_qubits,
                          const uint64_t *const *xmask, const int64_t *const *grp_off, const int32_t *n_groups,
                          const uint64_t *const *zmask, const double *const *coef, const int64_t *n_terms,
                          const uint64_t *const *sorted_states, const int64_t *n_states, int32_t split_bits,
                          rqx_stream_t stream) {
    g_err.clear();
    if (!out) return fail(RQX_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (n_problems < 1) return fail(RQX_ERR_ARG, "n_problems must be >= 1");
    if (!n_qubits || !xmask || !grp_off || !n_groups || !zmask || !coef || !n_terms || !sorted_states || !n_states)
        return fail(RQX_ERR_ARG, "NULL argument array");
    std::vector<ProblemIn> in;
    for (int k = 0; k < n_problems; ++k)
        in.push_back({n_qubits[k], xmask[k], grp_off[k], n_groups[k], zmask[k], coef[k], n_terms[k], sorted_states[k],
                      n_states[k]});
    return build_plan(out, in, -1, split_bits, 1, static_cast<hipStream_t>(stream));
}

int rqx_plan_create_device(rqx_plan **out, int32_t n_qubits, const uint64_t *xmask, const int64_t *grp_off,
                           int32_t n_groups, const uint64_t *zmask, const double *coef, int64_t n_terms,
                           const uint64_t *sorted_states, int64_t n_states, int64_t n_total, int32_t split_bits,
                           rqx_stream_t stream) {
    g_err.clear();
    if (!out) return fail(RQX_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (n_total < n_states) return fail(RQX_ERR_ARG, "negative size or n_total < n_states");
    std::vector<ProblemIn> in = {{n_qubits, xmask, grp_off, n_groups, zmask, coef, n_terms, sorted_states, n_states}};
    return build_plan(out, in, n_total, split_bits, 0, static_cast<hipStream_t>(stream), true);
}

int rqx_plan_digest(const rqx_plan *p, uint64_t *out, rqx_stream_t stream) {
    g_err.clear();
    if (!p || !out) return fail(RQX_ERR_ARG, "NULL argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_TRY(hipStreamSynchronize(st));
    const size_t N = static_cast<size_t>(p->n_states), M = static_cast<size_t>(p->n_segs),
                 C = static_cast<size_t>(p->n_classes), T = static_cast<size_t>(p->n_tiles);
    std::vector<uint32_t> hoff, hlog, soff(M + 1);
    size_t htab = 0;
    if (p->hashed && M) {
        hoff.resize(M);
        hlog.resize(M);
        HIP_TRY(hipMemcpy(hoff.data(), p->h_off, M * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(hlog.data(), p->h_log, M * 4, hipMemcpyDeviceToHost));
        htab = static_cast<size_t>(hoff[M - 1]) + (size_t(2) << hlog[M - 1]);
    }
    HIP_TRY(hipMemcpy(soff.data(), p->st_off, (M + 1) * 4, hipMemcpyDeviceToHost));
    const struct {
        const void *ptr;
        size_t bytes;
    } arrays[RQX_DIGEST_ARRAYS] = {
        {p->keys, N * 8}, {p->lo, N * 4}, {p->seg_hi, M * 8}, {p->seg_off, (M + 1) * 4},
        {p->cls_hi, C * 8}, {p->cls_goff, (C + 1) * 4}, {p->tile_row0, T * 4}, {p->tile_seg, T * 4},
        {p->tile_prob, T * 4}, {p->prob, p->probs.size() * sizeof(ProbDesc)},
        {p->h_off, hoff.size() * 4}, {p->h_log, hlog.size() * 4}, {p->h_tab, htab * 4},
        {p->st_off, (M + 1) * 4}, {p->st_list, static_cast<size_t>(soff[M]) * sizeof(TargetDesc)},
        {p->diag64, N * 16}};
    std::vector<uint64_t> buf;
    for (int a = 0; a < RQX_DIGEST_ARRAYS; ++a) {
        // FNV-1a over the array's 64-bit words (the tail zero-padded), seeded with its length
        buf.assign((arrays[a].bytes + 7) / 8, 0ull);
        if (arrays[a].bytes)
            HIP_TRY(hipMemcpy(buf.data(), arrays[a].ptr, arrays[a].bytes, hipMemcpyDeviceToHost));
        uint64_t h = 0xcbf29ce484222325ull ^ arrays[a].bytes;
        for (uint64_t w : buf) h = (h ^ w) * 0x100000001b3ull;
        out[a] = h;
    }
    return RQX_OK;
}

int rqx_pack_to_u64(const uint8_t *rows, int64_t n_rows, int32_t row_bytes, int32_t n_qubits, uint64_t *keys,
                    int64_t *n_real, rqx_stream_t stream) {
    g_err.clear();
    if (!n_real) return fail(RQX_ERR_ARG, "n_real is NULL");
    *n_real = 0;
    if (n_qubits < 1 || n_qubits > 64) return fail(RQX_ERR_ARG, "n_qubits must be in [1, 64]");
    if (row_bytes != (n_qubits + 8) / 8) return fail(RQX_ERR_ARG, "row_bytes must be ceil((n_qubits + 1) / 8)");
    if (n_rows < 0) return fail(RQX_ERR_ARG, "n_rows < 0");
    if (n_rows == 0) return RQX_OK;
    if (n_rows > 0x7fffffffLL) return fail(RQX_ERR_LIMIT, "n_rows must be < 2^31");
    if (!rows || !keys) return fail(RQX_ERR_ARG, "NULL argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int ni = static_cast<int>(n_rows);
    size_t t_sel = 0;
    HIP_TRY(hipcub::DeviceSelect::Flagged(nullptr, t_sel, static_cast<const uint64_t *>(nullptr),
                                          static_cast<const uint8_t *>(nullptr), keys, static_cast<int64_t *>(nullptr),
                                          ni, st));
    const size_t n = static_cast<size_t>(n_rows), o_all = align_up(t_sel), o_flag = o_all + align_up(n * 8),
                 o_num = o_flag + align_up(n);
    DevBuf b;
    HIP_TRY(b.alloc(o_num + 8));
    hipLaunchKernelGGL(k_unpack_rows, dim3(grid_for(n_rows)), dim3(ING_WG), 0, st, rows, n_rows, row_bytes, n_qubits,
                       b.at<uint64_t>(o_all), b.at<uint8_t>(o_flag));
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipcub::DeviceSelect::Flagged(b.p, t_sel, b.at<const uint64_t>(o_all), b.at<const uint8_t>(o_flag), keys,
                                          b.at<int64_t>(o_num), ni, st);
    int64_t num = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&num, b.at<int64_t>(o_num), 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(RQX_ERR_HIP, std::string("pack_to_u64: ") + hipGetErrorString(e));
    *n_real = num;
    return RQX_OK;
}

int rqx_plan_destroy(rqx_plan *plan) {
    g_err.clear();
    if (!plan) return RQX_OK;
    int prev = 0;
    HIP_TRY(hipGetDevice(&prev));
    HIP_TRY(hipSetDevice(plan->device));
    hipError_t e = hipFree(plan->block);
    const hipError_t es = hipFree(plan->st_block);
    if (e == hipSuccess) e = es;
    const hipError_t eh = hipFree(plan->hash_block);
    if (e == hipSuccess) e = eh;
    const hipError_t ea = hipFree(plan->al_block);
    if (e == hipSuccess) e = ea;
    const hipError_t ee = hipFree(plan->al_ent_block);
    if (e == hipSuccess) e = ee;
    const hipError_t ep = hipFree(plan->al_pblock);
    if (e == hipSuccess) e = ep;
    const hipError_t eq = hipFree(plan->al_pent_block);
    if (e == hipSuccess) e = eq;
    const hipError_t eb = hipFree(plan->bl_block);
    if (e == hipSuccess) e = eb;
    const hipError_t ebe = hipFree(plan->bl_ent_block);
    if (e == hipSuccess) e = ebe;
    for (HitCache &c : plan->cache) {
        const hipError_t ec = free_cache(c);
        if (e == hipSuccess) e = ec;
    }
    (void)hipSetDevice(prev);
    delete plan;
    if (e != hipSuccess) return fail(RQX_ERR_HIP, std::string("hipFree(plan): ") + hipGetErrorString(e));
    return RQX_OK;
}

int rqx_plan_get_info(const rqx_plan *p, rqx_plan_info *out) {
    g_err.clear();
    if (!p || !out) return fail(RQX_ERR_ARG, "NULL argument");
    const ProbHost &p0 = p->probs.front();
    out->n_qubits = p0.n_qubits;
    out->split_bits = p0.L;
    out->n_states = p->n_states;
    out->n_total = p->n_total;
    out->n_groups = p->n_groups;
    out->n_classes = p->n_classes;
    out->n_terms = p->n_terms;
    out->n_segments = p->n_segs;
    out->n_tiles = p->n_tiles;
    out->max_segment = p->max_seg;
    out->has_diagonal = p0.has_diag;
    out->is_real = p->is_real;
    int mg = 0;
    for (const ProbHost &ph : p->probs) mg = std::max(mg, ph.max_group_terms);
    out->max_group_terms = mg;
    out->n_problems = static_cast<int32_t>(p->probs.size());
    out->persistent = p->persistent;
    out->blocks_per_cu = p->blocks_per_cu > 0 ? p->blocks_per_cu : p->occupancy;
    out->n_cu = p->n_cu;
    return RQX_OK;
}

int rqx_plan_get_problem(const rqx_plan *p, int32_t k, rqx_problem_info *out) {
    g_err.clear();
    if (!p || !out) return fail(RQX_ERR_ARG, "NULL argument");
    if (k < 0 || k >= static_cast<int32_t>(p->probs.size())) return fail(RQX_ERR_ARG, "problem index out of range");
    const ProbHost &ph = p->probs[k];
    out->row_offset = ph.row_base;
    out->n_states = ph.n_states;
    out->term_offset = ph.term_base;
    out->n_terms = ph.n_terms;
    out->group_offset = ph.grp_base;
    out->n_groups = ph.n_groups;
    out->n_qubits = ph.n_qubits;
    out->split_bits = ph.L;
    out->n_segments = ph.n_segs;
    out->n_classes = ph.n_classes;
    out->n_tiles = ph.n_tiles;
    out->max_segment = ph.max_seg;
    out->has_diagonal = ph.has_diag;
    out->is_real = ph.is_real;
    out->cost = ph.cost;
    out->section_bits = ph.a_bits;
    out->alpha_rows = static_cast<int32_t>(ph.al_nlo);
    out->list_mode = ph.full ? 2 : ph.al_nlo ? 1 : 0;
    out->reserved = 0;
    return RQX_OK;
}

int rqx_plan_set_launch(rqx_plan *p, int32_t persistent, int32_t blocks_per_cu) {
    g_err.clear();
    if (!p) return fail(RQX_ERR_ARG, "plan is NULL");
    if (blocks_per_cu < 0 || blocks_per_cu > 64) return fail(RQX_ERR_ARG, "blocks_per_cu must be in [0, 64]");
    p->persistent = persistent ? 1 : 0;
    p->blocks_per_cu = blocks_per_cu;
    return RQX_OK;
}

int rqx_lookup(const rqx_plan *p, int32_t *idx, double *vals, rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (!idx && p->n_groups > 0 && p->n_total > 0) return fail(RQX_ERR_ARG, "idx is NULL");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t total = static_cast<int64_t>(p->n_groups) * p->n_total;
    if (total > 0) {
        HIP_TRY(fill_async(idx, int64_t(0), total, int32_t(-1), st));
        if (vals) HIP_TRY(fill_async(reinterpret_cast<double2 *>(vals), int64_t(0), total, make_double2(0.0, 0.0), st));
    }
    if (p->n_groups > 0) {
        KParams k = make_params(p, true);
        k.idx = idx;
        k.vals = reinterpret_cast<double2 *>(vals);
        HIP_TRY((launch_tiles<MODE_LOOKUP, double, true>(p, k, nullptr, nullptr, st)));
    }
    return RQX_OK;
}

int rqx_diagonal(const rqx_plan *p, double *diag, rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (!diag && p->n_total > 0) return fail(RQX_ERR_ARG, "diag is NULL");
    hipStream_t st = static_cast<hipStream_t>(stream);
    double2 *d = reinterpret_cast<double2 *>(diag);
    if (p->n_total > 0) HIP_TRY(fill_async(d, int64_t(0), p->n_total, make_double2(0.0, 0.0), st));
    if (p->n_states > 0)
        HIP_TRY(hipMemcpyAsync(d, p->diag64, static_cast<size_t>(p->n_states) * 16, hipMemcpyDeviceToDevice, st));
    return RQX_OK;
}

}  // extern "C"

namespace {
template <typename V, typename R>
hipError_t full_apply_launch(const rqx_plan *p, const void *coef, int64_t n_items, int64_t n_cols, const void *x,
                             void *y, hipStream_t st) {
    const int64_t want = (n_items + 255) / 256;
    const unsigned blocks = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(want, int64_t(p->n_cu) * 64)));
    if (p->is_real)
        hipLaunchKernelGGL((k_full_apply<V, R, true>), dim3(blocks), dim3(256), 0, st, p->g_x, p->g_toff, p->z, coef,
                           p->n_groups, n_items, n_cols, static_cast<const V *>(x), static_cast<V *>(y));
    else
        hipLaunchKernelGGL((k_full_apply<V, R, false>), dim3(blocks), dim3(256), 0, st, p->g_x, p->g_toff, p->z, coef,
                           p->n_groups, n_items, n_cols, static_cast<const V *>(x), static_cast<V *>(y));
    return hipGetLastError();
}
}  // namespace

extern "C" {

int rqx_full_apply(const rqx_plan *p, const void *x, void *y, int64_t n_cols, int32_t dtype, rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (p->probs.size() != 1) return fail(RQX_ERR_ARG, "rqx_full_apply needs a single-problem plan");
    const int nq = p->probs[0].n_qubits;
    if (nq > RQX_FULL_MAX_QUBITS) return fail(RQX_ERR_LIMIT, "rqx_full_apply: more than RQX_FULL_MAX_QUBITS qubits");
    if (n_cols < 1) return fail(RQX_ERR_ARG, "n_cols must be positive");
    if (!x || !y) return fail(RQX_ERR_ARG, "x or y is NULL");
    if (x == y) return fail(RQX_ERR_ARG, "x and y must not alias");
    if (int rc = check_dtype(p, dtype)) return rc;
    const int64_t dim = int64_t(1) << nq;
    if (dim > INT64_MAX / n_cols) return fail(RQX_ERR_LIMIT, "rqx_full_apply: 2^n x n_cols overflows");
    const int64_t n_items = dim * n_cols;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (p->n_groups == 0) {
        switch (dtype) {
            case RQX_F64: HIP_TRY(fill_async(static_cast<double *>(y), int64_t(0), n_items, 0.0, st)); break;
            case RQX_C128: HIP_TRY(fill_async(static_cast<double2 *>(y), int64_t(0), n_items, make_double2(0.0, 0.0), st)); break;
            case RQX_F32: HIP_TRY(fill_async(static_cast<float *>(y), int64_t(0), n_items, 0.0f, st)); break;
            default: HIP_TRY(fill_async(static_cast<float2 *>(y), int64_t(0), n_items, make_float2(0.f, 0.f), st)); break;
        }
        return RQX_OK;
    }
    switch (dtype) {
        case RQX_F64: HIP_TRY((full_apply_launch<double, double>(p, p->coef64, n_items, n_cols, x, y, st))); break;
        case RQX_C128: HIP_TRY((full_apply_launch<double2, double>(p, p->coef64, n_items, n_cols, x, y, st))); break;
        case RQX_F32: HIP_TRY((full_apply_launch<float, float>(p, p->coef32, n_items, n_cols, x, y, st))); break;
        default: HIP_TRY((full_apply_launch<float2, float>(p, p->coef32, n_items, n_cols, x, y, st))); break;
    }
    return RQX_OK;
}

// the matvec's kernels over the plan's rows: the cached kernel (a built cache of
// dtype's precision) or the matrix-free kernels
hipError_t matvec_kernels(const rqx_plan *p, const void *x, void *y, int32_t dtype, bool cached, hipStream_t st) {
    const int slot = is_fp64(dtype) ? 1 : 0;
    if (cached) {
        const HitCache &c = p->cache[slot];
        bool any_diag = false;
        for (const ProbHost &ph : p->probs) any_diag = any_diag || ph.has_diag;
        const void *diag = !any_diag ? nullptr
                         : slot == 1 ? static_cast<const void *>(p->diag64)
                                     : static_cast<const void *>(p->diag32);
        switch (dtype) {
            case RQX_F64: return spmv_launch<double, false>(p, c, diag, x, y, st);
            case RQX_C128: return spmv_launch<double, true>(p, c, diag, x, y, st);
            case RQX_F32: return spmv_launch<float, false>(p, c, diag, x, y, st);
            default: return spmv_launch<float, true>(p, c, diag, x, y, st);
        }
    }
    const KParams k = make_params(p, is_fp64(dtype));
    return launch_dtype<MODE_MATVEC>(p, k, x, y, dtype, st);
}

// After a cache build: for each complex dtype the new caches serve whose matvec
// would run the product kernel, both paths timed on zeros (one run after one
// warm-up each: a few matvecs; both give the same bits) and cache_slower set where
// the product kernel is faster.  Real dtypes keep the cache.
hipError_t choose_paths(rqx_plan *p, bool fp32, bool fp64, hipStream_t st) {
    std::vector<int32_t> dts;
    if (fp32 && product_plan(p, RQX_C64)) dts.push_back(RQX_C64);
    if (fp64 && product_plan(p, RQX_C128)) dts.push_back(RQX_C128);
    if (dts.empty() || p->n_total == 0) return hipSuccess;
    void *buf = nullptr;
    const size_t bytes = static_cast<size_t>(p->n_total) * 16;
    hipError_t e = hipMalloc(&buf, 2 * bytes);
    if (e != hipSuccess) return e;
    hipEvent_t ev[2] = {nullptr, nullptr};
    e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    if (e == hipSuccess) e = fill_async(static_cast<uint32_t *>(buf), int64_t(0), static_cast<int64_t>(bytes / 4), 0u, st);
    void *x = buf, *y = static_cast<char *>(buf) + bytes;
    for (size_t i = 0; i < dts.size() && e == hipSuccess; ++i) {
        float ms[2] = {0.f, 0.f};
        for (int cached = 0; cached < 2 && e == hipSuccess; ++cached) {
            e = matvec_kernels(p, x, y, dts[i], cached == 1, st);
            if (e == hipSuccess) e = hipEventRecord(ev[0], st);
            if (e == hipSuccess) e = matvec_kernels(p, x, y, dts[i], cached == 1, st);
            if (e == hipSuccess) e = hipEventRecord(ev[1], st);
            if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms[cached], ev[0], ev[1]);
        }
        if (e == hipSuccess) p->cache_slower[dts[i]] = ms[0] < ms[1] ? 1 : 0;
    }
    for (hipEvent_t v : ev)
        if (v) (void)hipEventDestroy(v);
    const hipError_t ef = hipFree(buf);
    return e != hipSuccess ? e : ef;
}

int rqx_matvec(const rqx_plan *p, const void *x, void *y, int32_t dtype, rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (p->n_total == 0) return RQX_OK;
    if (!x || !y) return fail(RQX_ERR_ARG, "x or y is NULL");
    if (x == y) return fail(RQX_ERR_ARG, "x and y must not alias");
    if (int rc = check_dtype(p, dtype)) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const bool cached = p->use_cache && p->cache[is_fp64(dtype) ? 1 : 0].built && p->n_states > 0 &&
                        !(p->use_cache == 2 && p->cache_slower[dtype]);
    HIP_TRY(matvec_kernels(p, x, y, dtype, cached, st));
    if (p->n_total > p->n_states) {
        const int64_t a = p->n_states, b = p->n_total;
        switch (dtype) {
            case RQX_F64: HIP_TRY(fill_async(static_cast<double *>(y), a, b, 0.0, st)); break;
            case RQX_C128: HIP_TRY(fill_async(static_cast<double2 *>(y), a, b, make_double2(0.0, 0.0), st)); break;
            case RQX_F32: HIP_TRY(fill_async(static_cast<float *>(y), a, b, 0.f, st)); break;
            default: HIP_TRY(fill_async(static_cast<float2 *>(y), a, b, make_float2(0.f, 0.f), st)); break;
        }
    }
    return RQX_OK;
}

int rqx_count(const rqx_plan *p, int64_t *counts, rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (!counts) return fail(RQX_ERR_ARG, "counts is NULL");
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_fill<int64_t>, dim3(1), dim3(64), 0, st, counts, int64_t(0), int64_t(3), int64_t(0));
    HIP_TRY(hipGetLastError());
    KParams k = make_params(p, true);
    k.counts = reinterpret_cast<unsigned long long *>(counts);
    HIP_TRY((launch_tiles<MODE_COUNT, double, true>(p, k, nullptr, nullptr, st)));
    return RQX_OK;
}

int rqx_plan_hits(const rqx_plan *p, int64_t *rows, int32_t *cols, double *vals, int64_t capacity, int64_t *count,
                  rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (!count) return fail(RQX_ERR_ARG, "count is NULL");
    if (capacity < 0 || (capacity > 0 && (!rows || !cols || !vals))) return fail(RQX_ERR_ARG, "bad output arrays");
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_TRY(fill_async(count, int64_t(0), int64_t(1), int64_t(0), st));
    KParams k = make_params(p, true);
    k.coo_count = reinterpret_cast<unsigned long long *>(count);
    k.coo_row = rows;
    k.coo_col = cols;
    k.coo_val = reinterpret_cast<double2 *>(vals);
    k.coo_capacity = capacity;
    HIP_TRY((launch_tiles<MODE_COO, double, true>(p, k, nullptr, nullptr, st)));
    return RQX_OK;
}

int rqx_grad_coef(const rqx_plan *p, const void *x, const void *gy, double *grad, int32_t dtype,
                  rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (p->n_terms == 0) return RQX_OK;
    if (!grad) return fail(RQX_ERR_ARG, "grad is NULL");
    if (p->n_states > 0 && (!x || !gy)) return fail(RQX_ERR_ARG, "x or gy is NULL");
    if (int rc = check_dtype(p, dtype)) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_TRY(fill_async(grad, int64_t(0), p->n_terms, 0.0, st));
    KParams k = make_params(p, is_fp64(dtype));
    k.gy = gy;
    k.grad = grad;
    HIP_TRY(launch_dtype<MODE_GRAD>(p, k, x, nullptr, dtype, st));
    return RQX_OK;
}

int rqx_expect_terms(const rqx_plan *p, const void *psi, double *out, int32_t dtype, rqx_stream_t stream) {
    return rqx_grad_coef(p, psi, psi, out, dtype, stream);
}

int rqx_plan_update_coef(rqx_plan *p, const double *coef, int32_t is_real, rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (p->n_terms == 0) return RQX_OK;
    if (!coef) return fail(RQX_ERR_ARG, "coef is NULL");
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_TRY(hipMemcpyAsync(p->coef64, coef, static_cast<size_t>(p->n_terms) * 16, hipMemcpyDeviceToDevice, st));
    const unsigned tb = static_cast<unsigned>((p->n_terms + 255) / 256);
    hipLaunchKernelGGL(k_narrow, dim3(tb), dim3(256), 0, st, p->coef64, p->coef32, p->n_terms);
    HIP_TRY(hipGetLastError());
    launch_group_tables(p, st);
    HIP_TRY(hipGetLastError());
    if (p->al_ent_block) HIP_TRY(launch_alpha_lists(p, true, st));   // values from the new tables
    if (p->bl_ent_block) HIP_TRY(launch_beta_lists(p, true, st));
    for (const ProbHost &ph : p->probs) {
        if (ph.n_states == 0) continue;
        const unsigned blocks = static_cast<unsigned>((ph.n_states + 255) / 256);
        const uint32_t k0 = static_cast<uint32_t>(ph.term_base);
        hipLaunchKernelGGL(k_diag, dim3(blocks), dim3(256), 0, st, p->keys + ph.row_base, ph.n_states, p->z, p->coef64,
                           k0, k0 + ph.diag_terms, p->diag64 + ph.row_base, p->diag32 + ph.row_base);
        HIP_TRY(hipGetLastError());
    }
    p->is_real = is_real ? 1 : 0;
    // cached hits keep their columns; values are rewritten on the stream, or the
    // cache is released when the value type changes (real <-> complex)
    for (int slot = 1; slot >= 0; --slot) {
        HitCache &c = p->cache[slot];
        if (!c.built) continue;
        if (c.cplx != (p->is_real ? 0 : 1)) {
            HIP_TRY(hipStreamSynchronize(st));
            HIP_TRY(release_cache(p, slot));
        } else if (!(c.shared && slot == 0)) {      // a shared pair is refilled in one pass
            HIP_TRY(fill_values(p, slot, st));
        }
    }
    return RQX_OK;
}

int rqx_plan_build_cache(rqx_plan *p, int32_t precision, int64_t max_bytes, rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    const bool both = precision == 0;
    const int slot = both ? 1 : cache_slot(precision);
    if (slot < 0) return fail(RQX_ERR_ARG, "precision must be 0 (both), 32 or 64");
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(release_cache(p, slot));
    if (both) HIP_TRY(release_cache(p, 0));
    if (p->n_states == 0) return RQX_OK;
    const int64_t n = p->n_states;
    if (n > INT32_MAX) return fail(RQX_ERR_LIMIT, "hit cache columns are int32: at most 2^31 - 1 rows");
    const int64_t n_slices = (n + SLICE - 1) / SLICE;

    // section B's SELL order (kept with the cache): positions, permutation, chunk widths
    const size_t ns64 = align_up(static_cast<size_t>(n_slices) * SLICE);
    const int64_t n_wg = (n + SPMV_WG - 1) / SPMV_WG;
    const size_t sell_bytes = 2 * ns64 + align_up(static_cast<size_t>(n_slices) * NCH * 2) +
                              align_up(static_cast<size_t>(n_wg) * 8) + 256;
    void *sell = nullptr;
    HIP_TRY(hipMalloc(&sell, sell_bytes));
    uint8_t *bpos = static_cast<uint8_t *>(sell), *bperm = bpos + ns64;
    uint16_t *bw = reinterpret_cast<uint16_t *>(bperm + ns64);
    int2 *win = reinterpret_cast<int2 *>(reinterpret_cast<char *>(bw) + align_up(static_cast<size_t>(n_slices) * NCH * 2));
    int *win_max_dev = reinterpret_cast<int *>(reinterpret_cast<char *>(win) + align_up(static_cast<size_t>(n_wg) * 8));
    int win_max = 0;
    // pass 1: non-diagonal hits per row (section B at [0, n), section A at
    // [n, 2n)), then each slice's section widths and SELL order (after 2n: A
    // widths, B widths, B chunk-width sums, a spare word per slice, the hit totals)
    uint32_t *rowcnt = nullptr;
    hipError_t e = hipMalloc(&rowcnt, static_cast<size_t>(2 * n + 4 * n_slices + 4) * 4);
    if (e != hipSuccess) {
        (void)hipFree(sell);
        return fail(RQX_ERR_HIP, std::string("cache row count: ") + hipGetErrorString(e));
    }
    KParams k = make_params(p, true);
    k.rowcnt = rowcnt;
    k.rowcnt_a = rowcnt + n;
    uint32_t *w_dev = rowcnt + 2 * n;
    unsigned long long *nnz_dev = reinterpret_cast<unsigned long long *>(w_dev + 4 * n_slices);
    e = fill_async<unsigned long long>(nnz_dev, 0, 2, 0ull, st);
    if (e == hipSuccess) e = fill_async<int>(win_max_dev, 0, 1, 0, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_stage_windows, dim3(static_cast<unsigned>((n_wg + 255) / 256)), dim3(256), 0, st, p->keys,
                           p->prob_rows, p->prob, static_cast<int>(p->probs.size()), n, n_wg, win, win_max_dev);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&win_max, win_max_dev, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = launch_tiles<MODE_ROWCOUNT, double, true>(p, k, nullptr, nullptr, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_slice_widths, dim3(static_cast<unsigned>((n_slices * SLICE + 255) / 256)), dim3(256), 0, st,
                           rowcnt, rowcnt + n, n, n_slices, w_dev, w_dev + n_slices, w_dev + 2 * n_slices, bpos, bperm,
                           bw, nnz_dev);
        e = hipGetLastError();
    }
    std::vector<uint32_t> wab(static_cast<size_t>(4 * n_slices + 4));
    if (e == hipSuccess) e = hipMemcpyAsync(wab.data(), w_dev, wab.size() * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    const hipError_t ef = hipFree(rowcnt);
    if (e == hipSuccess) e = ef;
    if (e != hipSuccess) {
        (void)hipFree(sell);
        return fail(RQX_ERR_HIP, std::string("cache row count: ") + hipGetErrorString(e));
    }

    // offsets of the column pass's full layout and of the stored section B
    const std::vector<uint32_t> wa(wab.begin(), wab.begin() + n_slices), wb(wab.begin() + n_slices, wab.begin() + 2 * n_slices),
        bsum(wab.begin() + 2 * n_slices, wab.begin() + 3 * n_slices);
    int64_t nnz = 0, nnz_b = 0;
    std::memcpy(&nnz, wab.data() + 4 * n_slices, 8);
    std::memcpy(&nnz_b, wab.data() + 4 * n_slices + 2, 8);
    std::vector<int64_t> off(static_cast<size_t>(n_slices), 0), aoff(static_cast<size_t>(n_slices), 0),
        boff(static_cast<size_t>(n_slices), 0);
    int64_t full = 0, a_levels = 0, b_entries = 0;
    for (int64_t s = 0; s < n_slices; ++s) {
        const uint32_t w_a = wa[static_cast<size_t>(s)], w_b = wb[static_cast<size_t>(s)];
        off[static_cast<size_t>(s)] = full;
        aoff[static_cast<size_t>(s)] = a_levels;
        boff[static_cast<size_t>(s)] = b_entries;
        full += static_cast<int64_t>(w_a + w_b) * SLICE;
        a_levels += w_a;
        b_entries += static_cast<int64_t>(bsum[static_cast<size_t>(s)]) * (CH * PACK);
    }
    const int cplx = p->is_real ? 0 : 1;
    const size_t vb = cache_val_bytes(slot, cplx);
    const size_t vb_lo = both ? cache_val_bytes(0, cplx) : 0;
    const size_t ns8 = align_up(static_cast<size_t>(n_slices) * 8), ns4 = align_up(static_cast<size_t>(n_slices) * 4);
    const size_t o_valb = 0, o_valb_lo = align_up(static_cast<size_t>(b_entries) * vb),
                 o_aval = o_valb_lo + (both ? align_up(static_cast<size_t>(b_entries) * vb_lo) : 0),
                 o_aval_lo = o_aval + align_up(static_cast<size_t>(a_levels) * vb),
                 o_boff = o_aval_lo + (both ? align_up(static_cast<size_t>(a_levels) * vb_lo) : 0),
                 o_avoff = o_boff + ns8, o_coff = o_avoff + ns8, o_aoff = o_coff + ns8,
                 o_wa = o_aoff + ns8, o_wb = o_wa + ns4, o_adel = o_wb + ns4,
                 o_acode = o_adel + align_up(static_cast<size_t>(a_levels) * 4),
                 total = o_acode + align_up(static_cast<size_t>(a_levels) * 4);
    // scratch of the column pass (columns of the full layout, section-A groups,
    // slice offsets, per-slice counts); the stored columns and section-A values
    // allocated later are at most as many as the full layout holds
    const size_t nf = static_cast<size_t>(std::max<int64_t>(full, 1));
    const size_t na = static_cast<size_t>(std::max<int64_t>(a_levels, 1)) * SLICE;
    const size_t scratch = align_up(nf * 4) + align_up(na * 4) + ns8 * 2;
    const size_t stored_max = align_up(nf * 4) + align_up(na * (vb + vb_lo));
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    const int64_t limit = max_bytes > 0 ? max_bytes : static_cast<int64_t>(free_b / 10 * 9);
    const int64_t need = static_cast<int64_t>(total + scratch + stored_max);
    if (need > limit) {
        (void)hipFree(sell);
        return fail(RQX_ERR_LIMIT, "hit cache needs " + std::to_string(need) + " bytes (while building), limit " +
                                       std::to_string(limit));
    }

    HitCache c;
    c.sell_block = sell;
    c.bpos = bpos;
    c.bperm = bperm;
    c.bw = bw;
    c.win = win;
    c.win_max = win_max;
    e = hipMalloc(&c.block, total);
    if (e != hipSuccess) {
        (void)free_cache(c);
        return fail(RQX_ERR_HIP, std::string("cache allocation: ") + hipGetErrorString(e));
    }
    char *b = static_cast<char *>(c.block);
    c.valb = b + o_valb;
    c.aval = b + o_aval;
    c.boff = reinterpret_cast<int64_t *>(b + o_boff);
    c.avoff = reinterpret_cast<int64_t *>(b + o_avoff);
    c.coff = reinterpret_cast<int64_t *>(b + o_coff);
    c.aoff = reinterpret_cast<int64_t *>(b + o_aoff);
    c.wa = reinterpret_cast<uint32_t *>(b + o_wa);
    c.wb = reinterpret_cast<uint32_t *>(b + o_wb);
    c.adel = reinterpret_cast<int32_t *>(b + o_adel);
    c.acode = reinterpret_cast<int32_t *>(b + o_acode);
    c.full_entries = full;
    c.nnz = nnz;
    c.n_slices = n_slices;
    c.a_levels = a_levels;
    c.valb_entries = b_entries;
    c.b_hits = nnz_b;
    c.cplx = cplx;
    c.shared = both;
    int32_t *col_full = nullptr, *gid_a = nullptr;
    int64_t *off_d = nullptr;
    uint32_t *counts = nullptr;
    auto release_scratch = [&]() {
        for (void *q : {static_cast<void *>(col_full), static_cast<void *>(gid_a), static_cast<void *>(off_d),
                        static_cast<void *>(counts)})
            (void)hipFree(q);
    };
    auto fail_build = [&](hipError_t err, const char *what) {
        (void)hipStreamSynchronize(st);
        release_scratch();
        p->cache[slot] = HitCache{};
        if (both) p->cache[0] = HitCache{};
        (void)free_cache(c);
        return fail(RQX_ERR_HIP, std::string(what) + hipGetErrorString(err));
    };
    e = hipMalloc(&col_full, nf * 4);
    if (e == hipSuccess) e = hipMalloc(&gid_a, na * 4);
    if (e == hipSuccess) e = hipMalloc(&off_d, ns8);
    if (e == hipSuccess) e = hipMalloc(&counts, ns8 + 3 * 8);
    unsigned long long *pad_dev = counts ? reinterpret_cast<unsigned long long *>(
                                               reinterpret_cast<char *>(counts) + ns8) : nullptr;
    if (e == hipSuccess) e = fill_async<unsigned long long>(pad_dev, 0, 3, 0ull, st);
    if (e == hipSuccess) e = hipMemcpyAsync(off_d, off.data(), off.size() * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(c.boff, boff.data(), boff.size() * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(c.aoff, aoff.data(), aoff.size() * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(c.wa, wa.data(), wa.size() * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(c.wb, wb.data(), wb.size() * 4, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return fail_build(e, "cache allocation: ");
    // pass 2: columns of every hit, groups of the section-A hits, section-B values
    e = fill_columns(p, slot, c, off_d, col_full, gid_a, both ? static_cast<void *>(b + o_valb_lo) : nullptr, st);
    if (e != hipSuccess) return fail_build(e, "cache columns: ");
    // section-A level codes: offsets and single values
    const unsigned wave_blocks = static_cast<unsigned>((n_slices * SLICE + 255) / 256);
    hipLaunchKernelGGL(k_cache_levels, dim3(wave_blocks), dim3(256), 0, st, col_full, gid_a, off_d, c.wa, c.aoff,
                       p->keys, p->g_zor, p->g_tab, p->g_toff, n, n_slices, c.adel, c.acode, counts, counts + n_slices,
                       pad_dev);
    e = hipGetLastError();
    std::vector<uint32_t> hc(static_cast<size_t>(2 * n_slices));
    unsigned long long pads[3] = {0, 0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(hc.data(), counts, hc.size() * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(pads, pad_dev, sizeof(pads), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail_build(e, "cache levels: ");
    std::vector<int64_t> coff(static_cast<size_t>(n_slices), 0), avoff(static_cast<size_t>(n_slices), 0);
    int64_t col_entries = 0, a_entries = 0;
    for (int64_t s = 0; s < n_slices; ++s) {
        coff[static_cast<size_t>(s)] = col_entries;
        avoff[static_cast<size_t>(s)] = a_entries;
        col_entries += static_cast<int64_t>(hc[static_cast<size_t>(s)]) * SLICE +
                       static_cast<int64_t>(bsum[static_cast<size_t>(s)]) * (CH * PACK);
        a_entries += static_cast<int64_t>(hc[static_cast<size_t>(n_slices + s)]) * SLICE;
    }
    const size_t o_vala_lo = align_up(static_cast<size_t>(std::max<int64_t>(a_entries, 1)) * vb);
    e = hipMalloc(&c.col_block, align_up(static_cast<size_t>(std::max<int64_t>(col_entries, 1)) * 4));
    if (e == hipSuccess)
        e = hipMalloc(&c.vala_block, o_vala_lo + (both ? static_cast<size_t>(std::max<int64_t>(a_entries, 1)) * vb_lo : 0));
    if (e != hipSuccess) return fail_build(e, "cache allocation: ");
    c.col = static_cast<int32_t *>(c.col_block);
    c.vala = c.vala_block;
    c.col_entries = col_entries;
    c.vala_entries = a_entries;
    c.pad_vala = static_cast<int64_t>(pads[0]);
    c.pad_cola = static_cast<int64_t>(pads[1]);
    c.pad_levels = static_cast<int64_t>(pads[2]);
    c.entries = b_entries + a_entries + a_levels;
    c.bytes = static_cast<int64_t>(total + align_up(static_cast<size_t>(col_entries) * 4) + o_vala_lo +
                                   (both ? static_cast<size_t>(a_entries) * vb_lo : 0) + sell_bytes);
    HitCache lo;                // the FP32 cache of a pair: its own values, everything else shared
    if (both) {
        lo = c;
        lo.block = nullptr;
        lo.col_block = nullptr;
        lo.vala_block = nullptr;
        lo.sell_block = nullptr;
        lo.valb = b + o_valb_lo;
        lo.aval = b + o_aval_lo;
        lo.vala = static_cast<char *>(c.vala_block) + o_vala_lo;
        lo.bytes = (b_entries + a_entries + a_levels) * static_cast<int64_t>(vb_lo);
    }
    p->cache[slot] = c;
    if (both) p->cache[0] = lo;
    e = hipMemcpyAsync(c.coff, coff.data(), coff.size() * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(c.avoff, avoff.data(), avoff.size() * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_cache_columns, dim3(wave_blocks), dim3(256), 0, st, col_full, off_d, c.wa, c.aoff,
                           c.adel, c.bpos, c.bw, c.coff, n, n_slices, c.col);
        e = hipGetLastError();
    }
    // pass 3: the section-A values, straight into the stored layout
    if (e == hipSuccess) e = fill_values(p, slot, st, true);
    if (e == hipSuccess) e = hipStreamSynchronize(st);   // host vectors and the scratch arrays are released below
    if (e != hipSuccess) return fail_build(e, "cache columns and values: ");
    release_scratch();
    p->cache[slot].built = true;
    if (both) p->cache[0].built = true;
    e = choose_paths(p, slot == 0 || both, slot == 1, st);
    if (e != hipSuccess) return fail(RQX_ERR_HIP, std::string("cache path choice: ") + hipGetErrorString(e));
    return RQX_OK;
}

int rqx_plan_drop_cache(rqx_plan *p, int32_t precision) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (precision != 0 && cache_slot(precision) < 0) return fail(RQX_ERR_ARG, "precision must be 0, 32 or 64");
    HIP_TRY(hipDeviceSynchronize());
    for (int slot = 1; slot >= 0; --slot)
        if (p->cache[slot].built && (precision == 0 || cache_slot(precision) == slot)) HIP_TRY(release_cache(p, slot));
    return RQX_OK;
}

int rqx_plan_cache_info(const rqx_plan *p, int32_t precision, int64_t *out) {
    g_err.clear();
    if (!p || !out) return fail(RQX_ERR_ARG, "NULL argument");
    const int slot = cache_slot(precision);
    if (slot < 0) return fail(RQX_ERR_ARG, "precision must be 32 or 64");
    const HitCache &c = p->cache[slot];
    out[0] = c.built ? 1 : 0;
    out[1] = c.nnz;
    out[2] = c.entries;
    out[3] = c.bytes;
    out[4] = c.n_slices;
    out[5] = c.cplx;
    out[6] = c.col_entries;
    out[7] = c.a_levels;
    out[8] = c.valb_entries;
    out[9] = c.vala_entries;
    out[10] = c.b_hits;
    out[11] = c.pad_vala;
    out[12] = c.pad_cola;
    out[13] = c.pad_levels;
    return RQX_OK;
}

int rqx_plan_use_cache(rqx_plan *p, int32_t enable) {
    g_err.clear();
    if (!p) return fail(RQX_ERR_ARG, "plan is NULL");
    if (enable < 0 || enable > 2) return fail(RQX_ERR_ARG, "enable must be 0, 1 or 2");
    p->use_cache = enable;
    return RQX_OK;
}

int rqx_plan_matvec_path(const rqx_plan *p, int32_t dtype, int32_t *path) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (!path) return fail(RQX_ERR_ARG, "path is NULL");
    if (dtype < RQX_F64 || dtype > RQX_C64) return fail(RQX_ERR_ARG, "unknown dtype");
    const bool cached = p->use_cache && p->cache[is_fp64(dtype) ? 1 : 0].built && p->n_states > 0 &&
                        !(p->use_cache == 2 && p->cache_slower[dtype]);
    *path = cached ? 2 : product_plan(p, dtype) ? 1 : 0;
    return RQX_OK;
}


}  // extern "C"

namespace {
unsigned batch_chunks(const rqx_plan *p) { return static_cast<unsigned>(p->chunks); }

template <typename V, int NP>
void dots_launch(const rqx_plan *p, const PtrPack &pk, double *out, hipStream_t st) {
    hipLaunchKernelGGL((k_batch_dots<V, NP>), dim3(batch_chunks(p), p->probs.size()), dim3(256), 0, st, pk,
                       p->prob_rows, p->dot_part);
    hipLaunchKernelGGL(k_dots_finish, dim3(p->probs.size()), dim3(64), 0, st, p->dot_part,
                       static_cast<int>(batch_chunks(p)), NP, out);
}

template <typename V>
int dots_dispatch(const rqx_plan *p, int np, const PtrPack &pk, double *out, hipStream_t st) {
    switch (np) {
        case 1: dots_launch<V, 1>(p, pk, out, st); break;
        case 2: dots_launch<V, 2>(p, pk, out, st); break;
        case 3: dots_launch<V, 3>(p, pk, out, st); break;
        case 4: dots_launch<V, 4>(p, pk, out, st); break;
        case 5: dots_launch<V, 5>(p, pk, out, st); break;
        case 6: dots_launch<V, 6>(p, pk, out, st); break;
        case 7: dots_launch<V, 7>(p, pk, out, st); break;
        case 8: dots_launch<V, 8>(p, pk, out, st); break;
        default: dots_launch<V, 9>(p, pk, out, st); break;
    }
    HIP_TRY(hipGetLastError());
    return RQX_OK;
}

template <typename V, int NV>
void combine_launch(const rqx_plan *p, const PtrPack &pk, const double *coef, void *out, hipStream_t st) {
    hipLaunchKernelGGL((k_batch_combine<V, NV>), dim3(batch_chunks(p), p->probs.size()), dim3(256), 0, st, pk,
                       p->prob_rows, coef, out);
}

template <typename V>
int combine_dispatch(const rqx_plan *p, int nv, const PtrPack &pk, const double *coef, void *out, hipStream_t st) {
    switch (nv) {
        case 1: combine_launch<V, 1>(p, pk, coef, out, st); break;
        case 2: combine_launch<V, 2>(p, pk, coef, out, st); break;
        case 3: combine_launch<V, 3>(p, pk, coef, out, st); break;
        default: combine_launch<V, 4>(p, pk, coef, out, st); break;
    }
    HIP_TRY(hipGetLastError());
    return RQX_OK;
}
}  // namespace

extern "C" {

int rqx_batch_dots(const rqx_plan *p, int32_t n_pairs, const void *const *a, const void *const *b, double *out,
                   int32_t dtype, rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (n_pairs < 1 || n_pairs > MAXV) return fail(RQX_ERR_ARG, "n_pairs must be in [1, 9]");
    if (!a || !b || !out) return fail(RQX_ERR_ARG, "NULL argument");
    if (dtype < RQX_F64 || dtype > RQX_C64) return fail(RQX_ERR_ARG, "unknown dtype");
    PtrPack pk{};
    for (int q = 0; q < n_pairs; ++q) {
        if (!a[q] || !b[q]) return fail(RQX_ERR_ARG, "NULL vector pointer");
        pk.a[q] = a[q];
        pk.b[q] = b[q];
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    switch (dtype) {
        case RQX_F64: return dots_dispatch<double>(p, n_pairs, pk, out, st);
        case RQX_C128: return dots_dispatch<double2>(p, n_pairs, pk, out, st);
        case RQX_F32: return dots_dispatch<float>(p, n_pairs, pk, out, st);
        default: return dots_dispatch<float2>(p, n_pairs, pk, out, st);
    }
}

int rqx_batch_combine(const rqx_plan *p, int32_t n_in, const void *const *v, const double *coef, void *out,
                      int32_t dtype, rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (n_in < 1 || n_in > 4) return fail(RQX_ERR_ARG, "n_in must be in [1, 4]");
    if (!v || !coef || !out) return fail(RQX_ERR_ARG, "NULL argument");
    if (dtype < RQX_F64 || dtype > RQX_C64) return fail(RQX_ERR_ARG, "unknown dtype");
    PtrPack pk{};
    for (int q = 0; q < n_in; ++q) {
        if (!v[q]) return fail(RQX_ERR_ARG, "NULL vector pointer");
        pk.a[q] = v[q];
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    switch (dtype) {
        case RQX_F64: return combine_dispatch<double>(p, n_in, pk, coef, out, st);
        case RQX_C128: return combine_dispatch<double2>(p, n_in, pk, coef, out, st);
        case RQX_F32: return combine_dispatch<float>(p, n_in, pk, coef, out, st);
        default: return combine_dispatch<float2>(p, n_in, pk, coef, out, st);
    }
}

}  // extern "C"

namespace {
template <typename V, typename LV, int NIN, int NOUT, typename DL, bool LO_IN, bool LO_OUT>
hipError_t lpass(const rqx_plan *p, const LPass &a, unsigned chunks, hipStream_t st) {
    hipLaunchKernelGGL((k_locg_pass<V, LV, NIN, NOUT, DL, LO_IN, LO_OUT>), dim3(chunks, p->probs.size()), dim3(256),
                       0, st, a, p->prob_rows);
    return hipGetLastError();
}

LPass make_pass(const rqx_plan *p, double *work, int coef_slot, std::initializer_list<const void *> in,
                std::initializer_list<void *> out) {
    LPass a{};
    int q = 0;
    for (const void *v : in) a.in[q++] = v;
    q = 0;
    for (void *v : out) a.out[q++] = v;
    a.coef = work + 2 * coef_slot;
    a.part = p->dot_part;
    a.stride = LOCG_WORK;
    return a;
}

template <typename V> constexpr int32_t dtype_of();
template <> constexpr int32_t dtype_of<double>() { return RQX_F64; }
template <> constexpr int32_t dtype_of<double2>() { return RQX_C128; }
template <> constexpr int32_t dtype_of<float>() { return RQX_F32; }
template <> constexpr int32_t dtype_of<float2>() { return RQX_C64; }

// One fused LOCG step (see rqx_locg_step) for vector type V (low copy type LV).
template <typename V, typename LV>
int locg_step(const rqx_plan *p, bool low, bool real, void *const *vec, double *theta, double *rel, uint8_t *active,
              double *work, double tol, hipStream_t st) {
    void *x = vec[0], *ax = vec[1], *y = vec[2], *ay = vec[3], *pv = vec[4], *ap = vec[5], *plo = vec[6], *aplo = vec[7];
    const int K = static_cast<int>(p->probs.size());
    const unsigned chunks = batch_chunks(p);
    // stage s reduces the ndot dots of the pass before it into the slots from dot_slot on
    auto scalar = [&](int stage, int dot_slot, int ndot) {
        hipLaunchKernelGGL(k_locg_scalar, dim3(K), dim3(64), 0, st, stage, K, work, theta, rel, active, tol,
                           real ? 1 : 0, p->dot_part, static_cast<int>(chunks), dot_slot, ndot);
        return hipGetLastError();
    };
    // every coefficient slot is written by a stage and every dot slot by a reduction before
    // it is read, so the work space needs no clearing (its scale slot is the caller's)
    // 1: theta = <x, Ax>
    HIP_TRY((lpass<V, LV, 2, 0, Dots<0x01>, false, false>(p, make_pass(p, work, 0, {x, ax}, {}), chunks, st)));
    HIP_TRY(scalar(1, LW_D_XAX, 1));
    // 2: r = Ax - theta x (into p); <r, r>, <x, r>, <y, r>
    HIP_TRY((lpass<V, LV, 3, 1, Dots<0x33, 0x13, 0x23>, false, false>(
        p, make_pass(p, work, LW_C2, {ax, x, y}, {pv}), chunks, st)));
    HIP_TRY(scalar(2, LW_D_RR, 3));
    // 3: p = r - <x,r> x - <y,r> y; <x, p>, <y, p>, <p, p>
    HIP_TRY((lpass<V, LV, 3, 1, Dots<0x13, 0x23, 0x33>, false, false>(
        p, make_pass(p, work, LW_C3, {pv, x, y}, {pv}), chunks, st)));
    HIP_TRY(scalar(3, LW_D_XP, 3));
    // 4: second Gram-Schmidt pass, normalised (plus the low copy for the bulk matvec); Ap; 5: Rayleigh-Ritz dots
    if (low) {
        HIP_TRY((lpass<V, LV, 3, 2, Dots<>, false, true>(p, make_pass(p, work, LW_C4, {pv, x, y}, {pv, plo}), chunks,
                                                         st)));
        if (int rc = rqx_matvec(p, plo, aplo, dtype_of<LV>(), st)) return rc;
        HIP_TRY((lpass<V, LV, 6, 1, Dots<0x03, 0x04, 0x06, 0x13, 0x14, 0x16, 0x23, 0x24, 0x26>, true, false>(
            p, make_pass(p, work, LW_C5, {x, y, pv, ax, ay, aplo}, {ap}), chunks, st)));
    } else {
        HIP_TRY((lpass<V, LV, 3, 1, Dots<>, false, false>(p, make_pass(p, work, LW_C4, {pv, x, y}, {pv}), chunks, st)));
        if (int rc = rqx_matvec(p, pv, ap, dtype_of<V>(), st)) return rc;
        HIP_TRY((lpass<V, LV, 6, 0, Dots<0x03, 0x04, 0x05, 0x13, 0x14, 0x15, 0x23, 0x24, 0x25>, false, false>(
            p, make_pass(p, work, 0, {x, y, pv, ax, ay, ap}, {}), chunks, st)));
    }
    HIP_TRY(scalar(5, LW_D_RR9, 9));
    // 6: y' = c1 y + c2 p, Ay' likewise, x' = c0 x + y', Ax' likewise; <x', x'>, <x', y'>
    HIP_TRY((lpass<V, LV, 6, 4, Dots<0x88, 0x86>, false, false>(
        p, make_pass(p, work, LW_C6, {x, ax, y, ay, pv, ap}, {y, ay, x, ax}), chunks, st)));
    HIP_TRY(scalar(6, LW_D_XX, 2));
    // 7: x = x' / |x'|, Ax likewise; y = y' - <x, y'> x, Ay likewise; <y, y>
    HIP_TRY((lpass<V, LV, 4, 4, Dots<0x66>, false, false>(
        p, make_pass(p, work, LW_C7, {x, ax, y, ay}, {x, ax, y, ay}), chunks, st)));
    HIP_TRY(scalar(7, LW_D_YY2, 1));
    // 8: y and Ay normalised
    HIP_TRY((lpass<V, LV, 2, 2, Dots<>, false, false>(p, make_pass(p, work, LW_C8, {y, ay}, {y, ay}), chunks, st)));
    return RQX_OK;
}
}  // namespace

extern "C" {

int rqx_locg_step(const rqx_plan *p, int32_t dtype, int32_t low, void *const *vec, double *theta, double *rel,
                  uint8_t *active, double *work, double tol, rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (int rc = check_dtype(p, dtype)) return rc;
    if (low && dtype != RQX_F64 && dtype != RQX_C128)
        return fail(RQX_ERR_ARG, "low != 0 needs RQX_F64 or RQX_C128 vectors");
    if (!vec || !theta || !rel || !active || !work) return fail(RQX_ERR_ARG, "NULL argument");
    for (int q = 0; q < (low ? 8 : 6); ++q)
        if (!vec[q]) return fail(RQX_ERR_ARG, "NULL vector pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    switch (dtype) {
        case RQX_F64: return locg_step<double, float>(p, low != 0, true, vec, theta, rel, active, work, tol, st);
        case RQX_C128: return locg_step<double2, float2>(p, low != 0, false, vec, theta, rel, active, work, tol, st);
        case RQX_F32: return locg_step<float, float>(p, false, true, vec, theta, rel, active, work, tol, st);
        default: return locg_step<float2, float2>(p, false, false, vec, theta, rel, active, work, tol, st);
    }
}

}  // extern "C"

// ---------------------------------------------------------------- rqx_locg_step_static
// The work space of rqx_locg_step for callers that cannot allocate: a file-scope
// device array (each GPU holds its own copy until the process ends).  Host code
// uses its address from hipGetSymbolAddress, looked up once per device with that
// device current; the array's name is not a device address in host code.
static __device__ __align__(256) double scratch[RQX_LOCG_STATIC_PROBLEMS * LOCG_WORK];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch[MAX_DEVICES] = {};

// scratch's address on the current device (looked up on its first use there)
int scratch_on_current_device(double **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(RQX_ERR_LIMIT, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(RQX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<double *>(ptr);
    }
    *out = g_scratch[dev];
    return RQX_OK;
}

// work[k * LOCG_WORK + RQX_LOCG_SCALE] = scale[k] (0 without scale)
__global__ void k_locg_scale(double *__restrict__ work, const double *__restrict__ scale, int K) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < K) work[static_cast<int64_t>(k) * LOCG_WORK + RQX_LOCG_SCALE] = scale ? scale[k] : 0.0;
}
}  // namespace

extern "C" {

int rqx_locg_step_static(const rqx_plan *p, int32_t dtype, int32_t low, void *const *vec, double *theta,
                         double *rel, uint8_t *active, const double *scale, double tol, rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    const int K = static_cast<int>(p->probs.size());
    if (K > RQX_LOCG_STATIC_PROBLEMS)
        return fail(RQX_ERR_LIMIT, "rqx_locg_step_static: the plan has more than RQX_LOCG_STATIC_PROBLEMS problems");
    double *work = nullptr;
    if (int rc = scratch_on_current_device(&work)) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_locg_scale, dim3(static_cast<unsigned>((K + 63) / 64)), dim3(64), 0, st, work, scale, K);
    HIP_TRY(hipGetLastError());
    return rqx_locg_step(p, dtype, low, vec, theta, rel, active, work, tol, stream);
}

int rqx_eigenpair_batch(const double *mats, int32_t count, int32_t n, double *out, rqx_stream_t stream) {
    g_err.clear();
    if (n != 2 && n != 3) return fail(RQX_ERR_ARG, "n must be 2 or 3");
    if (count < 0) return fail(RQX_ERR_ARG, "count < 0");
    if (count == 0) return RQX_OK;
    if (!mats || !out) return fail(RQX_ERR_ARG, "NULL argument");
    hipLaunchKernelGGL(k_eigen_batch, dim3(static_cast<unsigned>((count + 63) / 64)), dim3(64), 0,
                       static_cast<hipStream_t>(stream), mats, count, n, out);
    HIP_TRY(hipGetLastError());
    return RQX_OK;
}

int rqx_eigenpair_small(const double *mat, int32_t n, double *out, rqx_stream_t stream) {
    g_err.clear();
    if (n != 2 && n != 3) return fail(RQX_ERR_ARG, "n must be 2 or 3");
    if (!mat || !out) return fail(RQX_ERR_ARG, "NULL argument");
    hipLaunchKernelGGL(k_eigen_small, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream), mat, n, out);
    HIP_TRY(hipGetLastError());
    return RQX_OK;
}

int rqx_eigenpair_small_host(const double *mat, int32_t n, double *out) {
    g_err.clear();
    if (n != 2 && n != 3) return fail(RQX_ERR_ARG, "n must be 2 or 3");
    if (!mat || !out) return fail(RQX_ERR_ARG, "NULL argument");
    eigen_small(mat, n, out);
    return RQX_OK;
}

int rqx_sort_unique_u64(const uint64_t *keys, int64_t n, uint64_t *out, int64_t *n_unique, rqx_stream_t stream) {
    g_err.clear();
    if (!n_unique) return fail(RQX_ERR_ARG, "n_unique is NULL");
    *n_unique = 0;
    if (n < 0) return fail(RQX_ERR_ARG, "n < 0");
    if (n == 0) return RQX_OK;
    if (n > 0x7fffffffLL) return fail(RQX_ERR_LIMIT, "n must be < 2^31");
    if (!keys || !out) return fail(RQX_ERR_ARG, "NULL argument");
    if (keys == out) return fail(RQX_ERR_ARG, "keys and out must not alias");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int ni = static_cast<int>(n);
    size_t t_sort = 0, t_uniq = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, t_sort, keys, out, ni, 0, 64, st));
    HIP_TRY(hipcub::DeviceSelect::Unique(nullptr, t_uniq, out, out, static_cast<int64_t *>(nullptr), ni, st));
    const size_t tmp_bytes = align_up(std::max(t_sort, t_uniq));
    void *ws = nullptr;
    HIP_TRY(hipMalloc(&ws, tmp_bytes + align_up(static_cast<size_t>(n) * 8) + 256));
    char *w = static_cast<char *>(ws);
    uint64_t *sorted = reinterpret_cast<uint64_t *>(w + tmp_bytes);
    int64_t *d_num = reinterpret_cast<int64_t *>(w + tmp_bytes + align_up(static_cast<size_t>(n) * 8));
    hipError_t e = hipcub::DeviceRadixSort::SortKeys(w, t_sort, keys, sorted, ni, 0, 64, st);
    if (e == hipSuccess) e = hipcub::DeviceSelect::Unique(w, t_uniq, sorted, out, d_num, ni, st);
    int64_t h_num = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&h_num, d_num, sizeof(int64_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    const hipError_t ef = hipFree(ws);
    if (e != hipSuccess) return fail(RQX_ERR_HIP, std::string("sort_unique: ") + hipGetErrorString(e));
    if (ef != hipSuccess) return fail(RQX_ERR_HIP, std::string("sort_unique free: ") + hipGetErrorString(ef));
    *n_unique = h_num;
    return RQX_OK;
}

}  // extern "C"
