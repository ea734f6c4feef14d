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
#include <string>
#include <type_traits>
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

enum { MODE_MATVEC = 0, MODE_LOOKUP = 1, MODE_COUNT = 2, MODE_GRAD = 3 };

struct KParams {
    const uint64_t *keys;
    const uint32_t *lo;
    const uint64_t *seg_hi;
    const uint32_t *seg_off;
    const uint64_t *cls_hi;
    const uint32_t *cls_goff;
    const uint32_t *g_xlo;
    const uint32_t *g_toff;
    const uint64_t *z;
    const void *coef;      // double2[T] or float2[T]
    const void *diag;      // double2[N] or float2[N]
    const uint32_t *tile_row0;
    const uint32_t *tile_seg;
    int32_t *idx;                  // lookup output
    double2 *vals;                 // lookup output (optional): matrix element C_j(s_i)
    unsigned long long *counts;    // count output
    const uint64_t *g_x;           // full X mask per group
    const void *gy;                // gradient mode: upstream gradient (same dtype as x)
    double *grad;                  // gradient mode: per-term output (accumulated)
    int64_t n_total;
    uint32_t lomask;
    int32_t n_classes;
    int32_t n_segs;
    int32_t has_diag;
};

template <typename R> struct C2;
template <> struct C2<double> { using T = double2; };
template <> struct C2<float> { using T = float2; };

template <typename R, bool CPLX> struct VT { using T = R; };
template <typename R> struct VT<R, true> { using T = typename C2<R>::T; };

template <typename R> __device__ inline R vzero(R *) { return R(0); }
__device__ inline double2 vzero(double2 *) { return make_double2(0.0, 0.0); }
__device__ inline float2 vzero(float2 *) { return make_float2(0.f, 0.f); }

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

// Map the hardware dispatch order (round robin over XCDs) to contiguous tile
// ranges per XCD, so neighbouring tiles share one XCD's L2.
__device__ inline uint32_t xcd_remap(uint32_t b, uint32_t nb) {
    const uint32_t xcd = b % NUM_XCD, local = b / NUM_XCD;
    const uint32_t q = nb / NUM_XCD, r = nb % NUM_XCD;
    return (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

// index of t in sorted a[0..n) or -1
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

// C_g(s) = sum_k coef_k (-1)^popcount(s & z_k)
template <typename R>
__device__ inline typename C2<R>::T group_coef(const KParams &p, uint32_t g, uint64_t s) {
    using CC = typename C2<R>::T;
    const CC *coef = static_cast<const CC *>(p.coef);
    CC acc;
    acc.x = R(0);
    acc.y = R(0);
    const uint32_t k1 = p.g_toff[g + 1];
    for (uint32_t k = p.g_toff[g]; k < k1; ++k) {
        const CC c = coef[k];
        const bool odd = __popcll(s & p.z[k]) & 1;
        acc.x += odd ? -c.x : c.x;
        acc.y += odd ? -c.y : c.y;
    }
    return acc;
}

// sum over the 64 lanes of a wave (every lane receives the total)
template <typename R>
__device__ inline R wave_sum(R v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

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
    const uint64_t xg = p.g_x[g];
    const uint32_t k1 = p.g_toff[g + 1];
    for (uint32_t k = p.g_toff[g]; k < k1; ++k) {
        const uint64_t zk = p.z[k];
        const int ph = __popcll(xg & zk) & 3;
        R val = hit ? re_phase(ph, wre, wim) : R(0);
        if (__popcll(s & zk) & 1) val = -val;
        val = wave_sum(val);
        if ((threadIdx.x & 63) == 0) atomicAdd(&p.grad[k], static_cast<double>(val));
    }
}

template <int MODE, typename R, bool CPLX>
__global__ __launch_bounds__(TILE) void k_tiles(const KParams p, const void *__restrict__ xin,
                                                void *__restrict__ yout) {
    using V = typename VT<R, CPLX>::T;
    using CC = typename C2<R>::T;
    extern __shared__ uint32_t s_keys[];      // staged target segment (max_segment keys)
    __shared__ int32_t s_tseg[TILE];

    const int tid = threadIdx.x;
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t seg = p.tile_seg[tile];
    const uint32_t i = p.tile_row0[tile] + tid;
    const bool active = i < p.seg_off[seg + 1];
    const uint64_t s = active ? p.keys[i] : 0ull;
    const uint32_t slo = static_cast<uint32_t>(s) & p.lomask;
    const uint64_t shi = p.seg_hi[seg];
    const V *x = static_cast<const V *>(xin);

    V acc = vzero(static_cast<V *>(nullptr));
    unsigned long long n_hit = 0, n_probe = 0;
    V gv = vzero(static_cast<V *>(nullptr));
    if constexpr (MODE == MODE_GRAD) {
        if (active) gv = static_cast<const V *>(p.gy)[i];
        if (p.has_diag) {
            R wre = R(0), wim = R(0);
            if (active) conj_mul(gv, x[i], wre, wim);
            grad_group<R>(p, 0u, s, active, wre, wim);
        }
    }
    if (p.has_diag && active && MODE != MODE_GRAD) {
        if constexpr (MODE == MODE_MATVEC) {
            const CC d = static_cast<const CC *>(p.diag)[i];
            fma_acc(acc, d, x[i]);
        } else if constexpr (MODE == MODE_LOOKUP) {
            p.idx[i] = static_cast<int32_t>(i);
            if (p.vals) p.vals[i] = static_cast<const double2 *>(p.diag)[i];
        } else {
            n_hit += 1;
        }
    }

    for (int cb = 0; cb < p.n_classes; cb += TILE) {
        const int c = cb + tid;
        int t = -1;
        if (c < p.n_classes) t = find_u64(p.seg_hi, p.n_segs, shi ^ p.cls_hi[c]);
        __syncthreads();              // readers of the previous chunk are done
        s_tseg[tid] = t;
        __syncthreads();
        const int cn = min(TILE, p.n_classes - cb);
        for (int cc = 0; cc < cn; ++cc) {
            const int ts = __builtin_amdgcn_readfirstlane(s_tseg[cc]);
            if (ts < 0) continue;
            const uint32_t ta = p.seg_off[ts];
            const uint32_t tn = p.seg_off[ts + 1] - ta;
            for (uint32_t k = tid; k < tn; k += TILE) s_keys[k] = p.lo[ta + k];
            __syncthreads();
            uint32_t g0 = p.cls_goff[cb + cc];
            const uint32_t g1 = p.cls_goff[cb + cc + 1];
            if (p.has_diag && g0 == 0) g0 = 1;
            auto handle = [&](uint32_t g, uint32_t base, uint32_t tl) {
                const bool hit = active && (s_keys[base] == tl);
                if constexpr (MODE == MODE_MATVEC) {
                    if (hit) fma_acc(acc, group_coef<R>(p, g, s), x[ta + base]);
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
                } else {
                    n_hit += hit;
                    n_probe += active;
                }
            };
            uint32_t g = g0;
            // four independent searches in lock step: the dependent LDS loads of
            // one search overlap with those of the other three
            for (; g + 4 <= g1; g += 4) {
                const uint32_t t0 = slo ^ p.g_xlo[g], t1 = slo ^ p.g_xlo[g + 1];
                const uint32_t t2 = slo ^ p.g_xlo[g + 2], t3 = slo ^ p.g_xlo[g + 3];
                uint32_t b0 = 0, b1 = 0, b2 = 0, b3 = 0, len = tn;
                while (len > 1) {
                    const uint32_t half = len >> 1;
                    b0 = (s_keys[b0 + half] <= t0) ? b0 + half : b0;
                    b1 = (s_keys[b1 + half] <= t1) ? b1 + half : b1;
                    b2 = (s_keys[b2 + half] <= t2) ? b2 + half : b2;
                    b3 = (s_keys[b3 + half] <= t3) ? b3 + half : b3;
                    len -= half;
                }
                handle(g, b0, t0);
                handle(g + 1, b1, t1);
                handle(g + 2, b2, t2);
                handle(g + 3, b3, t3);
            }
            for (; g < g1; ++g) {
                const uint32_t tl = slo ^ p.g_xlo[g];
                uint32_t base = 0, len = tn;
                while (len > 1) {
                    const uint32_t half = len >> 1;
                    base = (s_keys[base + half] <= tl) ? base + half : base;
                    len -= half;
                }
                handle(g, base, tl);
            }
            __syncthreads();          // before the next target overwrites s_keys
        }
    }

    if constexpr (MODE == MODE_MATVEC) {
        if (active) static_cast<V *>(yout)[i] = acc;
    } else if constexpr (MODE == MODE_COUNT) {
        // wave reduction, one atomic per wave
        for (int off = 32; off > 0; off >>= 1) {
            n_hit += __shfl_down(n_hit, off, 64);
            n_probe += __shfl_down(n_probe, off, 64);
        }
        if ((tid & 63) == 0) {
            atomicAdd(&p.counts[0], n_hit);
            atomicAdd(&p.counts[1], n_probe);
        }
        if (tid == 0) atomicAdd(&p.counts[2], static_cast<unsigned long long>(p.n_classes));
    }
}

template <typename V>
__global__ void k_fill(V *__restrict__ y, int64_t begin, int64_t end, V value) {
    const int64_t i = begin + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < end) y[i] = value;
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

void build_segments(Segmentation &sg, const uint64_t *keys, int64_t n, const uint64_t *xmask, int32_t J) {
    const int L = sg.L;
    sg.seg_hi.clear();
    sg.seg_off.clear();
    sg.cls_hi.clear();
    sg.cls_goff.clear();
    sg.max_seg = 0;
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t h = (L >= 64) ? 0 : keys[i] >> L;
        if (i == 0 || h != sg.seg_hi.back()) {
            sg.seg_hi.push_back(h);
            sg.seg_off.push_back(static_cast<uint32_t>(i));
        }
    }
    sg.seg_off.push_back(static_cast<uint32_t>(n));
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

// Estimated per-thread work of the tile kernel, from a deterministic sample of tiles.
double estimate_cost(const Segmentation &sg, bool has_diag) {
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
            cost += (tn + TILE - 1) / TILE + 12.0 + groups * (2.0 * ceil_log2(tn) + 4.0);
        }
        total += cost;
    }
    return total / n_sample * tile_seg.size();
}

}  // namespace

struct rqx_plan {
    int device = 0;
    int n_qubits = 0;
    int L = 0;
    int has_diag = 0;
    int is_real = 1;
    int max_seg = 0;
    int max_group_terms = 0;
    uint32_t diag_terms = 0;   // terms of the x == 0 group (group 0)
    int64_t n_states = 0, n_total = 0, n_terms = 0;
    int32_t n_groups = 0, n_classes = 0, n_segs = 0, n_tiles = 0;
    // device arrays (one allocation, carved)
    void *block = nullptr;
    uint64_t *keys = nullptr;
    uint32_t *lo = nullptr;
    uint64_t *seg_hi = nullptr;
    uint32_t *seg_off = nullptr;
    uint64_t *cls_hi = nullptr;
    uint32_t *cls_goff = nullptr;
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
};

namespace {

// dynamic LDS of the tile kernel: one staged segment of uint32 low keys
size_t seg_lds_bytes(const rqx_plan *p) { return static_cast<size_t>(std::max(p->max_seg, 1)) * 4; }

KParams make_params(const rqx_plan *p, bool fp64) {
    KParams k{};
    k.keys = p->keys;
    k.lo = p->lo;
    k.seg_hi = p->seg_hi;
    k.seg_off = p->seg_off;
    k.cls_hi = p->cls_hi;
    k.cls_goff = p->cls_goff;
    k.g_xlo = p->g_xlo;
    k.g_toff = p->g_toff;
    k.g_x = p->g_x;
    k.z = p->z;
    k.coef = fp64 ? static_cast<const void *>(p->coef64) : static_cast<const void *>(p->coef32);
    k.diag = fp64 ? static_cast<const void *>(p->diag64) : static_cast<const void *>(p->diag32);
    k.tile_row0 = p->tile_row0;
    k.tile_seg = p->tile_seg;
    k.n_total = p->n_total;
    k.lomask = (p->L >= 32) ? 0xffffffffu : ((1u << p->L) - 1u);
    k.n_classes = p->n_classes;
    k.n_segs = p->n_segs;
    k.has_diag = p->has_diag;
    return k;
}

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

int check_plan(const rqx_plan *p) {
    if (!p) return fail(RQX_ERR_ARG, "plan is NULL");
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != p->device)
        return fail(RQX_ERR_ARG, "plan was created on device " + std::to_string(p->device) +
                                     " but the current device is " + std::to_string(dev));
    return RQX_OK;
}

}  // namespace

extern "C" {

int rqx_version(void) { return 100; }

const char *rqx_last_error(void) { return g_err.c_str(); }

int rqx_plan_create(rqx_plan **out, int32_t n_qubits, const uint64_t *xmask, const int64_t *grp_off,
                    int32_t n_groups, const uint64_t *zmask, const double *coef, int64_t n_terms,
                    const uint64_t *sorted_states, int64_t n_states, int64_t n_total, int32_t split_bits,
                    rqx_stream_t stream) {
    g_err.clear();
    if (!out) return fail(RQX_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (n_qubits < 1 || n_qubits > 64) return fail(RQX_ERR_ARG, "n_qubits must be in [1, 64]");
    if (n_groups < 0 || n_terms < 0 || n_states < 0 || n_total < n_states)
        return fail(RQX_ERR_ARG, "negative size or n_total < n_states");
    if (n_total > 0x7fffffffLL) return fail(RQX_ERR_LIMIT, "n_total must be < 2^31");
    if (n_terms > 0x7fffffffLL) return fail(RQX_ERR_LIMIT, "n_terms must be < 2^31");
    if (n_groups > 0 && (!xmask || !grp_off)) return fail(RQX_ERR_ARG, "xmask/grp_off is NULL");
    if (n_terms > 0 && (!zmask || !coef)) return fail(RQX_ERR_ARG, "zmask/coef is NULL");
    if (n_states > 0 && !sorted_states) return fail(RQX_ERR_ARG, "sorted_states is NULL");
    const uint64_t qmask = (n_qubits == 64) ? ~0ull : ((1ull << n_qubits) - 1ull);
    if (n_groups > 0) {
        if (grp_off[0] != 0 || grp_off[n_groups] != n_terms)
            return fail(RQX_ERR_ARG, "grp_off must start at 0 and end at n_terms");
        for (int32_t j = 0; j < n_groups; ++j) {
            if (grp_off[j + 1] < grp_off[j]) return fail(RQX_ERR_ARG, "grp_off must be non-decreasing");
            if (xmask[j] & ~qmask) return fail(RQX_ERR_ARG, "xmask has bits beyond n_qubits");
            if (j > 0 && xmask[j] <= xmask[j - 1])
                return fail(RQX_ERR_ARG, "xmask must be strictly ascending (one group per X mask)");
        }
    } else if (n_terms != 0) {
        return fail(RQX_ERR_ARG, "terms given without groups");
    }
    for (int64_t k = 0; k < n_terms; ++k)
        if (zmask[k] & ~qmask) return fail(RQX_ERR_ARG, "zmask has bits beyond n_qubits");
    for (int64_t i = 0; i < n_states; ++i) {
        if (sorted_states[i] & ~qmask) return fail(RQX_ERR_ARG, "state has bits beyond n_qubits");
        if (i > 0 && sorted_states[i] <= sorted_states[i - 1])
            return fail(RQX_ERR_ARG, "states must be sorted ascending and unique");
    }

    rqx_plan *p = new rqx_plan();
    HIP_TRY(hipGetDevice(&p->device));
    p->n_qubits = n_qubits;
    p->n_states = n_states;
    p->n_total = n_total;
    p->n_terms = n_terms;
    p->n_groups = n_groups;
    p->has_diag = (n_groups > 0 && xmask[0] == 0) ? 1 : 0;
    for (int64_t k = 0; k < n_terms; ++k)
        if (coef[2 * k + 1] != 0.0) p->is_real = 0;
    for (int32_t j = 0; j < n_groups; ++j)
        p->max_group_terms = std::max<int>(p->max_group_terms, static_cast<int>(grp_off[j + 1] - grp_off[j]));
    p->diag_terms = p->has_diag ? static_cast<uint32_t>(grp_off[1]) : 0u;

    // ---- choose the split L
    std::vector<int8_t> hd(static_cast<size_t>(n_states), 64);
    for (int64_t i = 1; i < n_states; ++i) hd[i] = static_cast<int8_t>(highest_diff_bit(sorted_states[i - 1], sorted_states[i]));
    const int Lmax_bits = std::min(32, n_qubits);
    Segmentation best;
    if (split_bits >= 0) {
        if (split_bits > Lmax_bits) {
            delete p;
            return fail(RQX_ERR_ARG, "split_bits must be <= min(32, n_qubits)");
        }
        if (max_segment(hd, split_bits) > SEG_CAP) {
            delete p;
            return fail(RQX_ERR_LIMIT, "split_bits gives a segment larger than " + std::to_string(SEG_CAP) + " rows");
        }
        best.L = split_bits;
        build_segments(best, sorted_states, n_states, xmask, n_groups);
    } else {
        // largest feasible L (max segment is non-decreasing in L), then the
        // cheapest of the candidates below it by the sampled cost model
        int lo_l = 0, hi_l = Lmax_bits;
        while (lo_l < hi_l) {
            const int mid = (lo_l + hi_l + 1) / 2;
            if (max_segment(hd, mid) <= SEG_CAP) lo_l = mid; else hi_l = mid - 1;
        }
        const int Lfeas = lo_l;
        bool first = true;
        for (int L = Lfeas; L >= std::max(0, Lfeas - 8); --L) {
            Segmentation cand;
            cand.L = L;
            build_segments(cand, sorted_states, n_states, xmask, n_groups);
            cand.cost = estimate_cost(cand, p->has_diag);
            if (first || cand.cost < best.cost) {
                best = std::move(cand);
                first = false;
            }
        }
    }
    p->L = best.L;
    p->max_seg = best.max_seg;
    p->n_segs = static_cast<int32_t>(best.seg_hi.size());
    p->n_classes = static_cast<int32_t>(best.cls_hi.size());

    std::vector<uint32_t> tile_row0, tile_seg;
    for (int32_t m = 0; m < p->n_segs; ++m)
        for (uint32_t r = best.seg_off[m]; r < best.seg_off[m + 1]; r += TILE) {
            tile_row0.push_back(r);
            tile_seg.push_back(static_cast<uint32_t>(m));
        }
    p->n_tiles = static_cast<int32_t>(tile_row0.size());

    const uint32_t lomask = (p->L >= 32) ? 0xffffffffu : ((1u << p->L) - 1u);
    std::vector<uint32_t> lo(static_cast<size_t>(n_states));
    for (int64_t i = 0; i < n_states; ++i) lo[i] = static_cast<uint32_t>(sorted_states[i]) & lomask;
    std::vector<uint32_t> g_xlo(static_cast<size_t>(n_groups)), g_toff(static_cast<size_t>(n_groups) + 1);
    for (int32_t j = 0; j < n_groups; ++j) g_xlo[j] = static_cast<uint32_t>(xmask[j]) & lomask;
    for (int32_t j = 0; j <= n_groups; ++j) g_toff[j] = n_groups ? static_cast<uint32_t>(grp_off[j]) : 0u;
    if (n_groups == 0) g_toff.assign(1, 0u);
    std::vector<float2> c32(static_cast<size_t>(n_terms));
    for (int64_t k = 0; k < n_terms; ++k)
        c32[k] = make_float2(static_cast<float>(coef[2 * k]), static_cast<float>(coef[2 * k + 1]));

    // ---- one device block, carved
    const size_t n = static_cast<size_t>(n_states);
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t o = off;
        off += align_up(std::max<size_t>(bytes, 1));
        return o;
    };
    const size_t o_keys = carve(n * 8), o_lo = carve(n * 4), o_seghi = carve(best.seg_hi.size() * 8),
                 o_segoff = carve(best.seg_off.size() * 4), o_clshi = carve(best.cls_hi.size() * 8),
                 o_clsgoff = carve(std::max<size_t>(best.cls_goff.size(), 1) * 4), o_gxlo = carve(g_xlo.size() * 4),
                 o_gtoff = carve(g_toff.size() * 4), o_z = carve(static_cast<size_t>(n_terms) * 8),
                 o_c64 = carve(static_cast<size_t>(n_terms) * 16), o_c32 = carve(static_cast<size_t>(n_terms) * 8),
                 o_d64 = carve(n * 16), o_d32 = carve(n * 8), o_trow = carve(tile_row0.size() * 4),
                 o_tseg = carve(tile_seg.size() * 4), o_gx = carve(static_cast<size_t>(n_groups) * 8);
    hipError_t e = hipMalloc(&p->block, off);
    if (e != hipSuccess) {
        delete p;
        return fail(RQX_ERR_HIP, std::string("hipMalloc(plan): ") + hipGetErrorString(e));
    }
    char *b = static_cast<char *>(p->block);
    p->keys = reinterpret_cast<uint64_t *>(b + o_keys);
    p->lo = reinterpret_cast<uint32_t *>(b + o_lo);
    p->seg_hi = reinterpret_cast<uint64_t *>(b + o_seghi);
    p->seg_off = reinterpret_cast<uint32_t *>(b + o_segoff);
    p->cls_hi = reinterpret_cast<uint64_t *>(b + o_clshi);
    p->cls_goff = reinterpret_cast<uint32_t *>(b + o_clsgoff);
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

    hipStream_t st = static_cast<hipStream_t>(stream);
    auto up = [&](void *dst, const void *src, size_t bytes) -> hipError_t {
        if (bytes == 0) return hipSuccess;
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    };
    hipError_t errs[16] = {
        up(p->keys, sorted_states, n * 8),
        up(p->lo, lo.data(), n * 4),
        up(p->seg_hi, best.seg_hi.data(), best.seg_hi.size() * 8),
        up(p->seg_off, best.seg_off.data(), best.seg_off.size() * 4),
        up(p->cls_hi, best.cls_hi.data(), best.cls_hi.size() * 8),
        up(p->cls_goff, best.cls_goff.data(), best.cls_goff.size() * 4),
        up(p->g_xlo, g_xlo.data(), g_xlo.size() * 4),
        up(p->g_toff, g_toff.data(), g_toff.size() * 4),
        up(p->z, zmask, static_cast<size_t>(n_terms) * 8),
        up(p->coef64, coef, static_cast<size_t>(n_terms) * 16),
        up(p->coef32, c32.data(), static_cast<size_t>(n_terms) * 8),
        up(p->tile_row0, tile_row0.data(), tile_row0.size() * 4),
        up(p->tile_seg, tile_seg.data(), tile_seg.size() * 4),
        up(p->g_x, xmask, static_cast<size_t>(n_groups) * 8),
        hipSuccess, hipSuccess};
    for (hipError_t ei : errs)
        if (ei != hipSuccess) {
            (void)hipFree(p->block);
            delete p;
            return fail(RQX_ERR_HIP, std::string("hipMemcpyAsync(plan): ") + hipGetErrorString(ei));
        }
    if (n > 0) {
        const int threads = 256;
        const unsigned blocks = static_cast<unsigned>((n + threads - 1) / threads);
        const uint32_t k1 = p->has_diag ? static_cast<uint32_t>(grp_off[1]) : 0u;
        hipLaunchKernelGGL(k_diag, dim3(blocks), dim3(threads), 0, st, p->keys, static_cast<int64_t>(n), p->z,
                           p->coef64, 0u, k1, p->diag64, p->diag32);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            (void)hipFree(p->block);
            delete p;
            return fail(RQX_ERR_HIP, std::string("k_diag: ") + hipGetErrorString(e));
        }
    } else {
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            (void)hipFree(p->block);
            delete p;
            return fail(RQX_ERR_HIP, std::string("plan upload: ") + hipGetErrorString(e));
        }
    }
    *out = p;
    return RQX_OK;
}

int rqx_plan_destroy(rqx_plan *plan) {
    g_err.clear();
    if (!plan) return RQX_OK;
    int prev = 0;
    HIP_TRY(hipGetDevice(&prev));
    HIP_TRY(hipSetDevice(plan->device));
    const hipError_t e = hipFree(plan->block);
    (void)hipSetDevice(prev);
    delete plan;
    if (e != hipSuccess) return fail(RQX_ERR_HIP, std::string("hipFree(plan): ") + hipGetErrorString(e));
    return RQX_OK;
}

int rqx_plan_get_info(const rqx_plan *p, rqx_plan_info *out) {
    g_err.clear();
    if (!p || !out) return fail(RQX_ERR_ARG, "NULL argument");
    out->n_qubits = p->n_qubits;
    out->split_bits = p->L;
    out->n_states = p->n_states;
    out->n_total = p->n_total;
    out->n_groups = p->n_groups;
    out->n_classes = p->n_classes;
    out->n_terms = p->n_terms;
    out->n_segments = p->n_segs;
    out->n_tiles = p->n_tiles;
    out->max_segment = p->max_seg;
    out->has_diagonal = p->has_diag;
    out->is_real = p->is_real;
    out->max_group_terms = p->max_group_terms;
    return RQX_OK;
}

int rqx_lookup(const rqx_plan *p, int32_t *idx, double *vals, rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (!idx && p->n_groups > 0 && p->n_total > 0) return fail(RQX_ERR_ARG, "idx is NULL");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t total = static_cast<int64_t>(p->n_groups) * p->n_total;
    if (total > 0) {
        const unsigned blocks = static_cast<unsigned>((total + 255) / 256);
        hipLaunchKernelGGL(k_fill<int32_t>, dim3(blocks), dim3(256), 0, st, idx, int64_t(0), total, int32_t(-1));
        HIP_TRY(hipGetLastError());
        if (vals) {
            hipLaunchKernelGGL(k_fill<double2>, dim3(blocks), dim3(256), 0, st, reinterpret_cast<double2 *>(vals),
                               int64_t(0), total, make_double2(0.0, 0.0));
            HIP_TRY(hipGetLastError());
        }
    }
    if (p->n_tiles > 0 && p->n_groups > 0) {
        KParams k = make_params(p, true);
        k.idx = idx;
        k.vals = reinterpret_cast<double2 *>(vals);
        hipLaunchKernelGGL((k_tiles<MODE_LOOKUP, double, true>), dim3(p->n_tiles), dim3(TILE), seg_lds_bytes(p), st, k,
                           nullptr, nullptr);
        HIP_TRY(hipGetLastError());
    }
    return RQX_OK;
}

int rqx_diagonal(const rqx_plan *p, double *diag, rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (!diag && p->n_total > 0) return fail(RQX_ERR_ARG, "diag is NULL");
    hipStream_t st = static_cast<hipStream_t>(stream);
    double2 *d = reinterpret_cast<double2 *>(diag);
    if (p->n_total > 0) {
        const unsigned blocks = static_cast<unsigned>((p->n_total + 255) / 256);
        hipLaunchKernelGGL(k_fill<double2>, dim3(blocks), dim3(256), 0, st, d, int64_t(0), p->n_total,
                           make_double2(0.0, 0.0));
        HIP_TRY(hipGetLastError());
    }
    if (p->has_diag && p->n_states > 0)
        HIP_TRY(hipMemcpyAsync(d, p->diag64, static_cast<size_t>(p->n_states) * 16, hipMemcpyDeviceToDevice, st));
    return RQX_OK;
}

int rqx_matvec(const rqx_plan *p, const void *x, void *y, int32_t dtype, rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (p->n_total == 0) return RQX_OK;
    if (!x || !y) return fail(RQX_ERR_ARG, "x or y is NULL");
    if (x == y) return fail(RQX_ERR_ARG, "x and y must not alias");
    if (dtype < RQX_F64 || dtype > RQX_C64) return fail(RQX_ERR_ARG, "unknown dtype");
    const bool cplx = (dtype == RQX_C128 || dtype == RQX_C64);
    if (!cplx && !p->is_real)
        return fail(RQX_ERR_DTYPE, "real vector dtype requested but the Hamiltonian has complex coefficients");
    const bool fp64 = (dtype == RQX_F64 || dtype == RQX_C128);
    hipStream_t st = static_cast<hipStream_t>(stream);
    KParams k = make_params(p, fp64);
    if (p->n_tiles > 0) {
        const dim3 grid(p->n_tiles), block(TILE);
        const size_t lds = seg_lds_bytes(p);
        switch (dtype) {
            case RQX_F64: hipLaunchKernelGGL((k_tiles<MODE_MATVEC, double, false>), grid, block, lds, st, k, x, y); break;
            case RQX_C128: hipLaunchKernelGGL((k_tiles<MODE_MATVEC, double, true>), grid, block, lds, st, k, x, y); break;
            case RQX_F32: hipLaunchKernelGGL((k_tiles<MODE_MATVEC, float, false>), grid, block, lds, st, k, x, y); break;
            default: hipLaunchKernelGGL((k_tiles<MODE_MATVEC, float, true>), grid, block, lds, st, k, x, y); break;
        }
        HIP_TRY(hipGetLastError());
    }
    if (p->n_total > p->n_states) {
        const int64_t cnt = p->n_total - p->n_states;
        const unsigned blocks = static_cast<unsigned>((cnt + 255) / 256);
        switch (dtype) {
            case RQX_F64: hipLaunchKernelGGL(k_fill<double>, dim3(blocks), dim3(256), 0, st, static_cast<double *>(y), p->n_states, p->n_total, 0.0); break;
            case RQX_C128: hipLaunchKernelGGL(k_fill<double2>, dim3(blocks), dim3(256), 0, st, static_cast<double2 *>(y), p->n_states, p->n_total, make_double2(0.0, 0.0)); break;
            case RQX_F32: hipLaunchKernelGGL(k_fill<float>, dim3(blocks), dim3(256), 0, st, static_cast<float *>(y), p->n_states, p->n_total, 0.f); break;
            default: hipLaunchKernelGGL(k_fill<float2>, dim3(blocks), dim3(256), 0, st, static_cast<float2 *>(y), p->n_states, p->n_total, make_float2(0.f, 0.f)); break;
        }
        HIP_TRY(hipGetLastError());
    }
    return RQX_OK;
}

int rqx_count(const rqx_plan *p, int64_t *counts, rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (!counts) return fail(RQX_ERR_ARG, "counts is NULL");
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_TRY(hipMemsetAsync(counts, 0, 3 * sizeof(int64_t), st));
    if (p->n_tiles > 0) {
        KParams k = make_params(p, true);
        k.counts = reinterpret_cast<unsigned long long *>(counts);
        hipLaunchKernelGGL((k_tiles<MODE_COUNT, double, true>), dim3(p->n_tiles), dim3(TILE), seg_lds_bytes(p), st, k, nullptr,
                           nullptr);
        HIP_TRY(hipGetLastError());
    }
    return RQX_OK;
}

int rqx_grad_coef(const rqx_plan *p, const void *x, const void *gy, double *grad, int32_t dtype,
                  rqx_stream_t stream) {
    g_err.clear();
    if (int rc = check_plan(p)) return rc;
    if (p->n_terms == 0) return RQX_OK;
    if (!grad) return fail(RQX_ERR_ARG, "grad is NULL");
    if (p->n_states > 0 && (!x || !gy)) return fail(RQX_ERR_ARG, "x or gy is NULL");
    if (dtype < RQX_F64 || dtype > RQX_C64) return fail(RQX_ERR_ARG, "unknown dtype");
    const bool cplx = (dtype == RQX_C128 || dtype == RQX_C64);
    if (!cplx && !p->is_real)
        return fail(RQX_ERR_DTYPE, "real vector dtype requested but the Hamiltonian has complex coefficients");
    const bool fp64 = (dtype == RQX_F64 || dtype == RQX_C128);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const unsigned fb = static_cast<unsigned>((p->n_terms + 255) / 256);
    hipLaunchKernelGGL(k_fill<double>, dim3(fb), dim3(256), 0, st, grad, int64_t(0), p->n_terms, 0.0);
    HIP_TRY(hipGetLastError());
    if (p->n_tiles > 0) {
        KParams k = make_params(p, fp64);
        k.gy = gy;
        k.grad = grad;
        const dim3 grid(p->n_tiles), block(TILE);
        const size_t lds = seg_lds_bytes(p);
        switch (dtype) {
            case RQX_F64: hipLaunchKernelGGL((k_tiles<MODE_GRAD, double, false>), grid, block, lds, st, k, x, nullptr); break;
            case RQX_C128: hipLaunchKernelGGL((k_tiles<MODE_GRAD, double, true>), grid, block, lds, st, k, x, nullptr); break;
            case RQX_F32: hipLaunchKernelGGL((k_tiles<MODE_GRAD, float, false>), grid, block, lds, st, k, x, nullptr); break;
            default: hipLaunchKernelGGL((k_tiles<MODE_GRAD, float, true>), grid, block, lds, st, k, x, nullptr); break;
        }
        HIP_TRY(hipGetLastError());
    }
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
    if (p->n_states > 0) {
        const unsigned blocks = static_cast<unsigned>((p->n_states + 255) / 256);
        const uint32_t k1 = p->has_diag ? p->diag_terms : 0u;
        hipLaunchKernelGGL(k_diag, dim3(blocks), dim3(256), 0, st, p->keys, p->n_states, p->z, p->coef64, 0u, k1,
                           p->diag64, p->diag32);
        HIP_TRY(hipGetLastError());
    }
    p->is_real = is_real ? 1 : 0;
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
