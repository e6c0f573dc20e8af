// multirhs.hip -- H-matrix times a block of RR = 64 right-hand sides.
//
// Ensembles of fault realisations share one H-matrix, so
// the per-stage product becomes Y (N x RR) = H X.  The factors are read once
// for all right-hand sides, which moves the product from HBM-bound to
// arithmetic-bound.  X, Z, Y are row-major with RR values per row (RR = 64,
// unused columns zero).
//
//   phase 1  Zpart(item) = a1^T[:, t-range] X[t-range, :] for column chunks of
//            every low-rank block; zreduce adds the chunks of a block in
//            order -> Z (sum_kt x RR)
//   phase 2  per tile item: Ypart = sum over segments of A_seg (TR x ncol)
//            X_seg (ncol x RR), k-slices staged through LDS
// FP64 factors multiply on the FP64 matrix cores, in the packed layout of the
// single-vector product.  FP32 and BF16 factors multiply, in HBI_PREC_F32
// plans, on the FP32 matrix cores (twice the FP64 rate on this chip) and, in
// HBI_PREC_MIXED plans, on the BF16 matrix cores with every FP32 operand
// split into three BF16 parts (FP32-sized products; see the BF16-core
// section and mr_mode for how the two round).  In both, X is rounded once to
// an FP32 copy, Z is stored in FP32, FP32 sums of at most one (FP32 cores)
// or two (BF16 cores) 32-column slices are added into FP64 sums, and the
// phase-1 partials, the phase-2 partials and Y stay FP64.
#include "hmat.hpp"

#include <algorithm>
#include <cstdlib>
#include <map>
#include <vector>

namespace hbi {

constexpr int kRR = 64;      // right-hand sides per product
constexpr int kP1Rows = 64;     // a1^T rows per phase-1 item on the FP64 cores (one wave)
constexpr int kKS = 32;      // k-slice staged in LDS
constexpr int kMrThreads = 256;
constexpr int kKS32 = 32;         // columns per slice of the FP32-core phase-2 kernel
constexpr int kP1Rows32 = 64;     // a1^T rows per phase-1 item on the FP32 cores (two waves of 32 rows)
constexpr int kMaxCols32 = 4096;  // columns of one phase-2 item on the FP32 / BF16 cores at most (column map in LDS)

template <typename T> __device__ __forceinline__ double widen(T v);
template <> __device__ __forceinline__ double widen<double>(double v) { return v; }
template <> __device__ __forceinline__ double widen<float>(float v) { return (double)v; }
template <> __device__ __forceinline__ double widen<uint16_t>(uint16_t v) { return (double)__uint_as_float(((uint32_t)v) << 16); }

// ------------------------------------------------------------- phase 1 --
struct MItem1 {
    uint64_t off;      // byte offset of the first a1^T row
    int ldt, kt;       // row stride (elements), rows (<= the plan's cap: whole blocks of one column cluster, stacked)
    int t0, tn;        // column chunk [t0, t0+tn)
    int xoff;          // first x row of the column cluster
    int part;          // partial slot zpart[part][rows cap][RR], or -1: the item is the rows' whole
                       // column range and writes Z rows zrow[zr .. zr + kt) directly
    int dtype, zr;
    int r0;            // first row of the item within its partial slot (set per wave by the FP32-core kernel)
};
// FP32-core phase-2 item: the item's segments are consecutive in the packed
// layout, so they form one TR x ncol column-major block at byte offset off;
// column k multiplies X row cmap[cm0 + k] (Z row when bit 31 is set)
struct MItem2 {
    uint64_t off;
    int ncol, cm0, part, pad;
};
struct MBlk1 {         // rows roff .. roff+kt of partial slots [p0, p1) -> z rows zoff .. zoff+kt
    int p0, p1, kt, zoff, roff;
};

typedef double v4d __attribute__((ext_vector_type(4)));
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef uint32_t uvec4 __attribute__((ext_vector_type(4)));
typedef uint32_t uvec2 __attribute__((ext_vector_type(2)));
typedef float fvec2 __attribute__((ext_vector_type(2)));

// zeros read in place of masked factor values and X rows (never written)
static __device__ double g_zero_row[kRR];

// 2 consecutive factor values of one a1^T row (4, 8 or 16 bytes, aligned)
template <typename T> struct Run2;
template <> struct Run2<double> {
    dvec2 v;
    __device__ void load(const double *p) { v = __builtin_nontemporal_load((const dvec2 *)p); }
    __device__ double get(int e) const { return v[e]; }
};
template <> struct Run2<float> {
    fvec2 v;
    __device__ void load(const float *p) { v = __builtin_nontemporal_load((const fvec2 *)p); }
    __device__ double get(int e) const { return (double)v[e]; }
    __device__ float getf(int e) const { return v[e]; }
};
template <> struct Run2<uint16_t> {
    uint32_t v;
    __device__ void load(const uint16_t *p) { v = __builtin_nontemporal_load((const uint32_t *)p); }
    __device__ double get(int e) const { return (double)__uint_as_float(e ? (v & 0xffff0000u) : (v << 16)); }
    __device__ float getf(int e) const { return __uint_as_float(e ? (v & 0xffff0000u) : (v << 16)); }
};


// One wave per item: Zpart (kt x 64) = a1^T[rows, t0 .. t0+tn) X[t0 .. t0+tn, :]
// on the FP64 matrix cores (16x16x4 MFMA; lane l = li + 16 lk supplies
// A[li][lk] and B[lk][li] and holds D[lk + 4 q][li]).  The columns pass in
// 8-column steps k; the reduction index of MFMA s (0, 1) of step k is
// t = t0 + 8 k + 2 lk + s, so each lane reads 2 consecutive a1^T values of its
// row per row block, and MFMA column n of right-hand-side block j is
// right-hand side 4 n + j, so each lane reads 32 contiguous bytes of an X
// row.  No LDS: three register buffers keep the loads of the next two steps
// in flight while one step is multiplied; two waves per SIMD overlap too.
template <typename T, int NRB>
__device__ __forceinline__ void mr_p1_wave(const MItem1 &it, const uint8_t *__restrict__ buf,
                                           const double *__restrict__ X, double *__restrict__ zpart,
                                           const int *__restrict__ zrow, double *__restrict__ Z)
{
    const int lane = threadIdx.x % 64, li = lane % 16, lk = lane / 16;
    const T *a = (const T *)(buf + it.off);
    const int tend = it.t0 + it.tn;
    struct Buf {
        Run2<T> a[NRB];
        dvec2 x[2][2];
    };
    Buf b0, b1, b2;
    // a 2-run never crosses an item boundary (items start at multiples of 32)
    // and rows are zero-padded up to ldt (a multiple of 8), so one test covers A
    // Masked loads read the zero row instead: no branches and no selects on
    // loaded values, so the waits before a step's products cover that step only.
    auto load = [&](Buf &B, int k) {
        const int tb = it.t0 + 8 * k + 2 * lk;
        const bool ok = tb < tend;
#pragma unroll
        for (int rb = 0; rb < NRB; rb++) {
            const int row = 16 * rb + li;
            B.a[rb].load(ok && row < it.kt ? a + (size_t)row * it.ldt + tb : (const T *)g_zero_row);
        }
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const dvec2 *px = tb + s < tend ? (const dvec2 *)(X + (size_t)(it.xoff + tb + s) * kRR + 4 * li)
                                            : (const dvec2 *)g_zero_row;
            B.x[s][0] = px[0];
            B.x[s][1] = px[1];
        }
        __builtin_amdgcn_sched_barrier(0);   // keep each step's loads where they are issued
    };
    v4d acc[NRB][4];
#pragma unroll
    for (int rb = 0; rb < NRB; rb++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[rb][j] = v4d{0.0, 0.0, 0.0, 0.0};
    auto mul = [&](const Buf &B) {
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int rb = 0; rb < NRB; rb++) {
                const double av = B.a[rb].get(s);
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[rb][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, B.x[s][j >> 1][j & 1], acc[rb][j], 0, 0, 0);
            }
        __builtin_amdgcn_sched_barrier(0);
    };
    // steps past the item's end load zeros: nothing is guarded
    const int nk = (it.tn + 7) / 8;
    load(b0, 0);
    load(b1, 1);
    int k = 0;
    for (; k + 3 <= nk; k += 3) {
        load(b2, k + 2);
        mul(b0);
        load(b0, k + 3);
        mul(b1);
        load(b1, k + 4);
        mul(b2);
    }
    if (k < nk) mul(b0);
    if (k + 1 < nk) mul(b1);
#pragma unroll
    for (int rb = 0; rb < NRB; rb++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int row = 16 * rb + lk + 4 * q;
            if (row < it.kt) {
                double *d = it.part < 0 ? Z + (size_t)zrow[it.zr + row] * kRR
                                        : zpart + ((size_t)it.part * kP1Rows + row) * kRR;
                dvec2 *dst = (dvec2 *)(d + 4 * li);
                dst[0] = dvec2{acc[rb][0][q], acc[rb][1][q]};
                dst[1] = dvec2{acc[rb][2][q], acc[rb][3][q]};
            }
        }
}

// The same item on the FP32 matrix cores (16x16x4 FP32 MFMA: operands as in
// the FP64 form, result element q of lane l is D[4 (l / 16) + q][l % 16]).
// Each wave takes 32 rows of an item (NRB <= 2) so the FP64 sums fit beside
// the FP32 accumulators at two waves per SIMD; the two waves of an item read
// the same X rows close together (the second read hits the CU's cache).  A step multiplies half as long as an
// FP64 step, so a ring of NB register buffers keeps NB - 1 steps of loads in
// flight; the FP32 sums of every NB steps (8 NB columns) are added into the
// FP64 sums.
template <typename T, int NRB>
__device__ __forceinline__ void mr_p1_wave32(const MItem1 &it, const uint8_t *__restrict__ buf,
                                             const float *__restrict__ X, double *__restrict__ zpart,
                                             const int *__restrict__ zrow, float *__restrict__ Z)
{
    constexpr int prow = kP1Rows32;
    const int lane = threadIdx.x % 64, li = lane % 16, lk = lane / 16;
    const T *a = (const T *)(buf + it.off);
    const int tend = it.t0 + it.tn;
    struct Buf {
        Run2<T> a[NRB];
        fvec4 x[2];
    };
    constexpr int NB = 6;
    Buf bb[NB];
    auto load = [&](Buf &B, int k) {
        const int tb = it.t0 + 8 * k + 2 * lk;
        const bool ok = tb < tend;
#pragma unroll
        for (int rb = 0; rb < NRB; rb++) {
            const int row = 16 * rb + li;
            B.a[rb].load(ok && row < it.kt ? a + (size_t)row * it.ldt + tb : (const T *)g_zero_row);
        }
#pragma unroll
        for (int s = 0; s < 2; s++)
            B.x[s] = tb + s < tend ? *(const fvec4 *)(X + (size_t)(it.xoff + tb + s) * kRR + 4 * li)
                                   : *(const fvec4 *)g_zero_row;
        __builtin_amdgcn_sched_barrier(0);
    };
    fvec4 acc[NRB][4];
    v4d sum[NRB][4];
#pragma unroll
    for (int rb = 0; rb < NRB; rb++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            acc[rb][j] = fvec4{0.f, 0.f, 0.f, 0.f};
            sum[rb][j] = v4d{0.0, 0.0, 0.0, 0.0};
        }
    auto mul = [&](const Buf &B) {
#pragma unroll
        for (int s = 0; s < 2; s++) {
#pragma unroll
            for (int rb = 0; rb < NRB; rb++) {
                const float av = B.a[rb].getf(s);
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[rb][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, B.x[s][j], acc[rb][j], 0, 0, 0);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    auto flush = [&] {
#pragma unroll
        for (int rb = 0; rb < NRB; rb++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
#pragma unroll
                for (int q = 0; q < 4; q++) sum[rb][j][q] += (double)acc[rb][j][q];
                acc[rb][j] = fvec4{0.f, 0.f, 0.f, 0.f};
            }
    };
    const int nk = (it.tn + 7) / 8;
#pragma unroll
    for (int b = 0; b < NB - 1; b++) load(bb[b], b);
    int k = 0;
    for (; k + NB <= nk; k += NB) {
#pragma unroll
        for (int b = 0; b < NB; b++) {
            load(bb[(b + NB - 1) % NB], k + b + NB - 1);   // steps past the item's end load zeros
            mul(bb[b]);
        }
        flush();
    }
    // steps k .. nk - 1 (fewer than NB) are in bb[0 ..]
#pragma unroll
    for (int b = 0; b < NB - 1; b++)
        if (k + b < nk) mul(bb[b]);
    flush();
#pragma unroll
    for (int rb = 0; rb < NRB; rb++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int row = 16 * rb + 4 * lk + q;
            if (row < it.kt) {
                if (it.part < 0) {
                    *(fvec4 *)(Z + (size_t)zrow[it.zr + row] * kRR + 4 * li) =
                        fvec4{(float)sum[rb][0][q], (float)sum[rb][1][q], (float)sum[rb][2][q], (float)sum[rb][3][q]};
                } else {
                    dvec2 *dst = (dvec2 *)(zpart + ((size_t)it.part * prow + it.r0 + row) * kRR + 4 * li);
                    dst[0] = dvec2{sum[rb][0][q], sum[rb][1][q]};
                    dst[1] = dvec2{sum[rb][2][q], sum[rb][3][q]};
                }
            }
        }
}

__global__ __launch_bounds__(2 * kP1Rows32) void mr_lrvt32_kernel(const MItem1 *__restrict__ items,
                                                        const uint8_t *__restrict__ buf, const float *__restrict__ X,
                                                        double *__restrict__ zpart, const int *__restrict__ zrow,
                                                        float *__restrict__ Z, const int *__restrict__ gate)
{
    if (gate && *gate) return;
    MItem1 it = items[blockIdx.x];
    // wave w: rows 32 w .. 32 w + 31 of the item (none: the wave ends; no barrier follows)
    const int w = threadIdx.x / 64;
    if (it.kt <= 32 * w) return;
    it.kt = min(32, it.kt - 32 * w);
    it.r0 = 32 * w;
    it.zr += 32 * w;
    it.off += (uint64_t)32 * w * it.ldt * (it.dtype == kF32 ? 4 : 2);
    if (it.dtype == kF32) {
        if (it.kt > 16) mr_p1_wave32<float, 2>(it, buf, X, zpart, zrow, Z);
        else mr_p1_wave32<float, 1>(it, buf, X, zpart, zrow, Z);
    } else {
        if (it.kt > 16) mr_p1_wave32<uint16_t, 2>(it, buf, X, zpart, zrow, Z);
        else mr_p1_wave32<uint16_t, 1>(it, buf, X, zpart, zrow, Z);
    }
}

__global__ __launch_bounds__(64) void mr_lrvt_kernel(const MItem1 *__restrict__ items,
                                                     const uint8_t *__restrict__ buf, const double *__restrict__ X,
                                                     double *__restrict__ zpart, const int *__restrict__ zrow,
                                                     double *__restrict__ Z, const int *__restrict__ gate)
{
    if (gate && *gate) return;
    const MItem1 it = items[blockIdx.x];
    const int nrb = (it.kt + 15) / 16;
#define HBI_P1(TT)                                                  \
    switch (nrb) {                                                  \
    case 1: mr_p1_wave<TT, 1>(it, buf, X, zpart, zrow, Z); break;            \
    case 2: mr_p1_wave<TT, 2>(it, buf, X, zpart, zrow, Z); break;            \
    case 3: mr_p1_wave<TT, 3>(it, buf, X, zpart, zrow, Z); break;            \
    default: mr_p1_wave<TT, 4>(it, buf, X, zpart, zrow, Z); break;           \
    }
    if (it.dtype == kF64) { HBI_P1(double) }
    else if (it.dtype == kF32) { HBI_P1(float) }
    else { HBI_P1(uint16_t) }
#undef HBI_P1
}

// X rounded to the FP32 copy the FP32-core kernels read (4 values a thread)
__global__ void mr_narrow_kernel(size_t n4, const double *__restrict__ X, float *__restrict__ X32,
                                 const int *__restrict__ gate)
{
    if (gate && *gate) return;
    const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (q >= n4) return;
    const dvec2 a = ((const dvec2 *)X)[2 * q], b = ((const dvec2 *)X)[2 * q + 1];
    ((fvec4 *)X32)[q] = fvec4{(float)a[0], (float)a[1], (float)b[0], (float)b[1]};
}

// Z rows of the blocks whose phase-1 rows were split into column chunks:
// the partial slots summed in slot order.  Workgroup q takes rows
// r0 .. r0 + 3 of block blk (zq[q] = (blk, r0), rows < kt only), one
// right-hand side per thread; the slots are loaded eight at a time (one
// memory latency per eight instead of one per slot), then added in order.
template <typename OT>
__global__ __launch_bounds__(4 * kRR) void mr_zreduce_kernel(const MBlk1 *__restrict__ blks, const int2 *__restrict__ zq,
                                                            int prow, const double *__restrict__ zpart,
                                                            OT *__restrict__ Z, const int *__restrict__ gate)
{
    if (gate && *gate) return;
    const int2 q = zq[blockIdx.x];
    const MBlk1 B = blks[q.x];
    const int k = q.y + (int)threadIdx.x / kRR, rr = (int)threadIdx.x % kRR;
    if (k >= B.kt) return;
    constexpr int kLd = 8;
    const size_t stride = (size_t)prow * kRR;
    const double *src = zpart + ((size_t)B.p0 * prow + B.roff + k) * kRR + rr;
    const int np = B.p1 - B.p0;
    double s = 0.0;
    for (int p = 0; p < np; p += kLd) {
        double v[kLd];
#pragma unroll
        for (int u = 0; u < kLd; u++) v[u] = src[(size_t)min(p + u, np - 1) * stride];
#pragma unroll
        for (int u = 0; u < kLd; u++)
            if (p + u < np) s += v[u];
    }
    Z[(size_t)(B.zoff + k) * kRR + rr] = (OT)s;
}

// ------------------------------------------------------------- phase 2 --
// One workgroup (4 waves) per tile item: Ypart (TR x 64) = sum over the item's
// segments of A_seg (TR x ncol) X_seg (ncol x 64) on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64; FP32 / BF16 factors are widened exactly).  k-slices
// of KS columns of A and KS rows of X pass through registers (the next
// slice's loads overlap the current slice's MFMAs) into LDS; wave (wr, wc)
// owns rows [wr TR/2, (wr+1) TR/2) and right-hand sides [32 wc, 32 wc + 32).
// Operand layout of the 16x16x4 FP64 MFMA, lane l: A[l % 16][l / 16],
// B[l / 16][l % 16]; result element q of lane l is D[l / 16 + 4 q][l % 16]
// (rows interleaved by 4, unlike the FP32 form's D[4 (l / 16) + q]).

template <typename T, int TR>
__global__ __launch_bounds__(kMrThreads) void mr_tile_kernel(const Item2 *__restrict__ items, const Seg *__restrict__ segs,
                                                             const uint8_t *__restrict__ buf,
                                                             const double *__restrict__ X, const double *__restrict__ Z,
                                                             double *__restrict__ ypart, const int *__restrict__ gate)
{
    constexpr int KS = 16;
    constexpr int RB = TR / 32;          // 16-row blocks per wave
    constexpr int AP = TR + 16, XP = kRR + 16;   // padded LDS rows: the four k-rows of a fragment hit both bank halves
    constexpr int NA = KS * TR / kMrThreads, NX = KS * kRR / kMrThreads;
    __shared__ double as[KS][AP];
    __shared__ double xs[KS][XP];
    if (gate && *gate) return;
    const Item2 it = items[blockIdx.x];
    const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
    const int wr = w / 2, wc = w % 2;
    const int li = lane % 16, lk = lane / 16;
    T ra[NA];            // raw factor values: widened at the LDS store (see phase 1)
    double rx[NX];
    // slice c of segment g: columns c .. c+KS of the column-major segment are contiguous
    auto load = [&](const Seg &g, int c) {
        const T *a = (const T *)(buf + g.off);
        const double *xin = (g.src ? Z : X) + (size_t)g.xoff * kRR;
        const int kn = min(KS, g.ncol - c);
#pragma unroll
        for (int u = 0; u < NA; u++) {
            const int q = tid + u * kMrThreads;
            ra[u] = q / TR < kn ? a[(size_t)c * TR + q] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < NX; u++) {
            const int q = tid + u * kMrThreads;
            rx[u] = q / kRR < kn ? xin[(size_t)c * kRR + q] : 0.0;
        }
    };
    v4d acc[RB][2];
#pragma unroll
    for (int ib = 0; ib < RB; ib++)
#pragma unroll
        for (int jb = 0; jb < 2; jb++) acc[ib][jb] = v4d{0.0, 0.0, 0.0, 0.0};
    int s = it.seg0, c0 = 0;
    Seg sg{};
    if (s < it.seg1) {
        sg = segs[s];
        load(sg, 0);
    }
    while (s < it.seg1) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NA; u++) {
            const int q = tid + u * kMrThreads;
            as[q / TR][q % TR] = widen<T>(ra[u]);
        }
#pragma unroll
        for (int u = 0; u < NX; u++) {
            const int q = tid + u * kMrThreads;
            xs[q / kRR][q % kRR] = rx[u];
        }
        __syncthreads();
        c0 += KS;
        if (c0 >= sg.ncol) {
            s++;
            c0 = 0;
            if (s < it.seg1) sg = segs[s];
        }
        if (s < it.seg1) load(sg, c0);
#pragma unroll
        for (int kk = 0; kk < KS; kk += 4) {
            double bf[2];
#pragma unroll
            for (int jb = 0; jb < 2; jb++) bf[jb] = xs[kk + lk][32 * wc + 16 * jb + li];
#pragma unroll
            for (int ib = 0; ib < RB; ib++) {
                const double af = as[kk + lk][RB * 16 * wr + 16 * ib + li];
#pragma unroll
                for (int jb = 0; jb < 2; jb++)
                    acc[ib][jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(af, bf[jb], acc[ib][jb], 0, 0, 0);
            }
        }
    }
    double *o = ypart + (size_t)it.part * TR * kRR;
#pragma unroll
    for (int ib = 0; ib < RB; ib++)
#pragma unroll
        for (int jb = 0; jb < 2; jb++)
#pragma unroll
            for (int q = 0; q < 4; q++)
                o[(size_t)(RB * 16 * wr + 16 * ib + lk + 4 * q) * kRR + 32 * wc + 16 * jb + li] = acc[ib][jb][q];
}

// The same product on the FP32 matrix cores for FP32 / BF16 factors: 32-deep
// k-slices of FP32 values in LDS; the FP32 sums of every slice are added into
// FP64 sums.  An item is one contiguous column block (its segments back to
// back), so a slice is a plain 32-column stride of it whatever the segment
// widths (low-rank segments have k_t columns, often far fewer than 32); the
// X / Z row of each column comes from the item's column map, staged in LDS
// once.  Every thread moves a slice in 16-byte pieces through a ring of
// register stages; no branch surrounds a load, so the wait before a stage's
// LDS store covers that stage only and the next NS - 1 slices stay in flight.
// NWC waves share the right-hand sides (64 / NWC each) in each of the two
// row halves: 8 waves for 64-row tiles, two waves on every SIMD per item.
template <typename T, int TR, int NWC>
__global__ __launch_bounds__(2 * NWC * 64) void mr_tile32_kernel(const MItem2 *__restrict__ items,
                                                                 const int *__restrict__ cmap,
                                                                 const uint8_t *__restrict__ buf,
                                                                 const float *__restrict__ X,
                                                                 const float *__restrict__ Z,
                                                                 double *__restrict__ ypart, const int *__restrict__ gate)
{
    constexpr int KS = kKS32;
    constexpr int NT = 2 * NWC * 64;                          // threads
    constexpr int RB = TR / 32;                               // 16-row blocks per wave
    constexpr int CW = kRR / NWC;                             // right-hand sides per wave
    constexpr int JB = CW / 16;
    constexpr int AP = TR + 16, XP = kRR + 16;
    constexpr int EPC = 16 / (int)sizeof(T);                 // factor values per 16-byte piece
    constexpr int CA = KS * TR / EPC;                         // pieces of a factor slice
    constexpr int NA = (CA + NT - 1) / NT;
    constexpr int CX = KS * kRR / 4;                          // pieces of an X slice
    constexpr int NX = (CX + NT - 1) / NT;
    static_assert(JB >= 1, "slice split");
    // one spare row: pieces past a slice's end are stored there, so no store is guarded
    __shared__ __align__(16) float as[KS + 1][AP];
    __shared__ __align__(16) float xs[KS + 1][XP];
    __shared__ int cm[kMaxCols32];
    if (gate && *gate) return;
    const MItem2 it = items[blockIdx.x];
    const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
    const int wr = w / NWC, wc = w % NWC;
    const int li = lane % 16, lk = lane / 16;
    for (int k = tid; k < it.ncol; k += NT) cm[k] = cmap[it.cm0 + k];
    __syncthreads();
    const uint8_t *fb = buf + it.off;
    struct Stage {
        uvec4 a[NA];
        fvec4 x[NX];
    };
    auto store = [&](const Stage &S) {
#pragma unroll
        for (int u = 0; u < NA; u++) {
            const int q = tid + u * NT;
            const int e0 = q * EPC, col = q < CA ? e0 / TR : KS, row = q < CA ? e0 % TR : 0;
            if constexpr (sizeof(T) == 4) {
                *(fvec4 *)&as[col][row] = fvec4{__uint_as_float(S.a[u][0]), __uint_as_float(S.a[u][1]),
                                                __uint_as_float(S.a[u][2]), __uint_as_float(S.a[u][3])};
            } else {
#pragma unroll
                for (int h = 0; h < 2; h++)
                    *(fvec4 *)&as[col][row + 4 * h] =
                        fvec4{__uint_as_float(S.a[u][2 * h] << 16), __uint_as_float(S.a[u][2 * h] & 0xffff0000u),
                              __uint_as_float(S.a[u][2 * h + 1] << 16),
                              __uint_as_float(S.a[u][2 * h + 1] & 0xffff0000u)};
            }
        }
#pragma unroll
        for (int u = 0; u < NX; u++) {
            const int q = tid + u * NT;
            const int col = q < CX ? q * 4 / kRR : KS, e = q < CX ? q * 4 % kRR : 0;
            *(fvec4 *)&xs[col][e] = S.x[u];
        }
    };
    // loads of slice j (columns 32 j ..); past the item's end: the zero row
    auto gather = [&](Stage &S, int j) {
        const int c0 = j * KS;
#pragma unroll
        for (int u = 0; u < NA; u++) {
            const int q = tid + u * NT;
            const int col = c0 + q * EPC / TR, row = q * EPC % TR;
            const bool ok = q < CA && col < it.ncol;
            S.a[u] = *(ok ? (const uvec4 *)(fb + ((size_t)col * TR + row) * sizeof(T)) : (const uvec4 *)g_zero_row);
        }
#pragma unroll
        for (int u = 0; u < NX; u++) {
            const int q = tid + u * NT;
            const int col = c0 + q * 4 / kRR, e = q * 4 % kRR;
            const bool ok = q < CX && col < it.ncol;
            const int m = cm[ok ? col : 0];
            const float *src = m < 0 ? Z : X;
            S.x[u] = *(ok ? (const fvec4 *)(src + (size_t)(m & 0x7fffffff) * kRR + e) : (const fvec4 *)g_zero_row);
        }
    };
    fvec4 acc[RB][JB];
    v4d sum[RB][JB];
#pragma unroll
    for (int ib = 0; ib < RB; ib++)
#pragma unroll
        for (int jb = 0; jb < JB; jb++) {
            acc[ib][jb] = fvec4{0.f, 0.f, 0.f, 0.f};
            sum[ib][jb] = v4d{0.0, 0.0, 0.0, 0.0};
        }
    auto multiply = [&] {
#pragma unroll
        for (int kk = 0; kk < KS; kk += 4) {
            float bf[JB];
#pragma unroll
            for (int jb = 0; jb < JB; jb++) bf[jb] = xs[kk + lk][CW * wc + 16 * jb + li];
#pragma unroll
            for (int ib = 0; ib < RB; ib++) {
                const float af = as[kk + lk][RB * 16 * wr + 16 * ib + li];
#pragma unroll
                for (int jb = 0; jb < JB; jb++)
                    acc[ib][jb] = __builtin_amdgcn_mfma_f32_16x16x4f32(af, bf[jb], acc[ib][jb], 0, 0, 0);
            }
        }
#pragma unroll
        for (int ib = 0; ib < RB; ib++)
#pragma unroll
            for (int jb = 0; jb < JB; jb++) {
#pragma unroll
                for (int q = 0; q < 4; q++) sum[ib][jb][q] += (double)acc[ib][jb][q];
                acc[ib][jb] = fvec4{0.f, 0.f, 0.f, 0.f};
            }
    };
    // a ring of NS register stages: slice j sits in stage j % NS; storing it
    // frees the stage for slice j + NS
    constexpr int NS = 4;
    const int nsl = (it.ncol + KS - 1) / KS;
    Stage st[NS];
#pragma unroll
    for (int b = 0; b < NS; b++) gather(st[b], b);
    for (int i = 0; i < nsl; i += NS) {
#pragma unroll
        for (int b = 0; b < NS; b++) {
            __syncthreads();
            store(st[b]);
            __syncthreads();
            gather(st[b], i + b + NS);
            if (i + b < nsl) multiply();
        }
    }
    double *o = ypart + (size_t)it.part * TR * kRR;
#pragma unroll
    for (int ib = 0; ib < RB; ib++)
#pragma unroll
        for (int jb = 0; jb < JB; jb++)
#pragma unroll
            for (int q = 0; q < 4; q++)
                o[(size_t)(RB * 16 * wr + 16 * ib + 4 * lk + q) * kRR + CW * wc + 16 * jb + li] = sum[ib][jb][q];
}

__global__ void mr_tile_reduce_kernel(int n, int tr, const int *__restrict__ part0, const double *__restrict__ ypart,
                                      double *__restrict__ Y)
{
    const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (q >= (size_t)n * kRR) return;
    const int r = (int)(q / kRR), rr = (int)(q % kRR);
    const int t = r / tr, ro = r % tr;
    double s = 0.0;
    for (int p = part0[t]; p < part0[t + 1]; p++) s += ypart[((size_t)p * tr + ro) * kRR + rr];
    Y[q] = s;
}

// --------------------------------------------- the BF16 matrix cores --
// The same two phases for FP32 / BF16 factors on the BF16 matrix cores
// (v_mfma_f32_16x16x32_bf16).  Every FP32 operand value is split exactly
// into three BF16 parts, x = x_h + x_m + x_l (round to nearest at each step;
// the remainder is below 2^-24 |x|), and each 16 x 16 x 32 block product is
// the sum of the six part products with i + j <= 2 (the dropped ones are
// below 2^-24 |a| |x|): FP32-level products at six BF16 instructions, which
// run faster than the sixteen FP32 16x16x4 ones the same block needs.  The
// FP32 sums of every 32-column slice are added into FP64 sums, as on the FP32
// cores.
//
// A workgroup of NW waves takes an item of 16 NW rows; wave w owns rows
// 16 w .. 16 w + 15 and all 64 right-hand sides (four 16-column MFMA blocks).
// Per 32-column slice, the workgroup stages the slice's 32 X / Z rows once in
// LDS, already split: three planes of BF16 pairs (columns 2 kp, 2 kp + 1 of a
// right-hand side in one word), each right-hand side a row of 16 words plus 4
// of padding, the four 16-byte word groups of a row XOR-swizzled by the row's
// bits 4-5 -- so both the transposing stores (4 consecutive columns x 16
// right-hand-side quads per wave) and the 16-byte operand reads (16 rows per
// read phase) hit 64 distinct banks.  The A operand never passes through LDS:
// phase 2 reads a copy of its factors packed in MFMA lane order (lane l of a
// 16 x 32 block holds row l % 16, columns 8 (l / 16) .. + 7: 2 KB per block,
// two 16-byte loads per lane), phase 1 reads its a1^T rows directly (eight
// consecutive columns of a row are contiguous).  Two LDS buffers: the loads of
// slice s + 1 are issued before slice s is multiplied and stored after it,
// one barrier per slice.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
constexpr int kBfBufs = 2;                // LDS buffers of the B slices
constexpr int kBfStages = 2;              // register stages of the load ring (2: the fewest registers; with
                                          // more, narrower items in flight per CU beat deeper prefetch)
constexpr int kBfRW = 20;                 // LDS words per right-hand-side row of a B plane
constexpr int kBfPlane = kRR * kBfRW;     // words per plane
constexpr int kP1RowsBf = 64;             // a1^T rows per phase-1 item at most (4 waves; 128-row items on 8
                                          // waves read X half as often but measured slower: 0.076 vs 0.055 ms)

// phase-2 item of the BF16-core kernel: its factors in lane order from float
// index off16, columns through the column map as MItem2
struct MItem2b {
    uint64_t off16;
    int ncol, cm0, part, pad;
};

__device__ __forceinline__ uint32_t bf_pack(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(fvec2{a, b}, bf16x2));
}
__device__ __forceinline__ fvec2 bf_unpack(uint32_t w)
{
    return fvec2{__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u)};
}
// (a, b) -> three packed BF16 words (high, middle, low parts)
__device__ __forceinline__ void bf_split3(float a, float b, uint32_t &h, uint32_t &m, uint32_t &l)
{
#pragma clang fp contract(off)
    h = bf_pack(a, b);
    const fvec2 hf = bf_unpack(h);
    const float ra = a - hf[0], rb = b - hf[1];
    m = bf_pack(ra, rb);
    const fvec2 mf = bf_unpack(m);
    l = bf_pack(ra - mf[0], rb - mf[1]);
}
__device__ __forceinline__ bf16x8 as_bf8(uvec4 v) { return __builtin_bit_cast(bf16x8, v); }

template <int NW, bool P2, int NS, int NB, bool A16 = false>
__global__ __launch_bounds__(64 * NW) void mr_bf16_kernel(const MItem2b *__restrict__ items2,
                                                          const MItem1 *__restrict__ items1,
                                                          const int *__restrict__ cmap,
                                                          const float *__restrict__ a16,
                                                          const uint8_t *__restrict__ buf,
                                                          const float *__restrict__ X, const float *__restrict__ Z,
                                                          double *__restrict__ out, const int *__restrict__ zrow,
                                                          float *__restrict__ Zout, int prow,
                                                          const int *__restrict__ gate)
{
    constexpr int NT = 64 * NW;
    constexpr int PP = (256 + NT - 1) / NT;   // B pairs (two columns x four right-hand sides) per thread
    __shared__ __align__(16) uint32_t bs[NB][3][kBfPlane];
    extern __shared__ int cm[];   // phase 2: the item's column map (dynamic: the plan's widest item)
    if (gate && *gate) return;
    const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
    const int c = lane % 16, q = lane / 16;
    // Workgroups go to the 8 XCDs round-robin by block index, and each XCD has
    // its own L2: in phase 2, XCD x takes the x-th contiguous range of items
    // (neighbouring row tiles, which read mostly the same X / Z rows).  Phase-1
    // items are ordered largest first and keep the plain order (balance).
    const int nb = gridDim.x, qx = nb / 8, rx = nb % 8, x = blockIdx.x % 8;
    const int item = P2 ? x * qx + min(x, rx) + (int)blockIdx.x / 8 : (int)blockIdx.x;
    MItem2b it2{};
    MItem1 it1{};
    int ncol;
    if constexpr (P2) {
        it2 = items2[item];
        ncol = it2.ncol;
        for (int k = tid; k < ncol; k += NT) cm[k] = cmap[it2.cm0 + k];
        __syncthreads();
    } else {
        it1 = items1[item];
        ncol = it1.tn;
    }
    // phase 1: a wave whose 16 rows lie past the item's rows only stages B
    // (wave-uniform, so its A loads and MFMAs are skipped, not masked)
    const bool live = P2 || 16 * w < it1.kt;
    const int nsl = (ncol + 31) / 32;
    const float *zf = (const float *)g_zero_row;
    // input row of item column col (past the item's end: zeros)
    auto brow = [&](int col) -> const float * {
        if (col >= ncol) return zf;
        if constexpr (P2) {
            const int m = cm[col];
            return (m < 0 ? Z : X) + (size_t)(m & 0x7fffffff) * kRR;
        } else {
            return X + (size_t)(it1.xoff + it1.t0 + col) * kRR;
        }
    };
    // a ring of NS register stages: slice s is loaded into stage s % NS three
    // slices before it is multiplied; no branch surrounds a load (slices past
    // the end read zeros), so each wait covers one stage
    struct Stage {
        fvec4 b0[PP], b1[PP];   // B: columns 2 kp, 2 kp + 1 of right-hand sides 4 m .. 4 m + 3
        fvec4 a0, a1;           // A: row 16 w + c, columns 8 q .. 8 q + 7 of the slice
    };
    Stage st[NS];
    auto load = [&](Stage &S, int s) {
#pragma unroll
        for (int u = 0; u < PP; u++) {
            const int p = min(tid + u * NT, 255), kp = p / 16, m = p % 16;
            S.b0[u] = *(const fvec4 *)(brow(32 * s + 2 * kp) + 4 * m);
            S.b1[u] = *(const fvec4 *)(brow(32 * s + 2 * kp + 1) + 4 * m);
        }
        if constexpr (P2) {
            const float *ap = s < nsl ? a16 + it2.off16 + ((size_t)(s * NW + w) * 64 + lane) * 8 : zf;
            S.a0 = *(const fvec4 *)ap;
            S.a1 = *(const fvec4 *)(ap + 4);
        } else if (live) {
            // FP32 rows: 8 values in two 16-byte loads; BF16 rows (A16): one 16-byte load (read twice)
            const int row = 16 * w + c, col = it1.t0 + 32 * s + 8 * q;
            const bool ok = row < it1.kt && col < it1.t0 + it1.tn;
            const uint8_t *ap = ok ? buf + it1.off + ((size_t)row * it1.ldt + col) * (A16 ? 2 : 4)
                                   : (const uint8_t *)zf;
            const fvec4 v0 = *(const fvec4 *)ap, v1 = *(const fvec4 *)(ap + (A16 ? 0 : 16));
            S.a0 = v0;
            S.a1 = v1;
        }
    };
    auto widen_a = [&](const Stage &S, fvec4 &a0, fvec4 &a1) {
        a0 = S.a0;
        a1 = S.a1;
        if constexpr (!P2 && A16) {   // BF16 factors: the 16 bytes hold 8 values
            const uvec4 v = __builtin_bit_cast(uvec4, S.a0);
            a0 = fvec4{__uint_as_float(v[0] << 16), __uint_as_float(v[0] & 0xffff0000u),
                       __uint_as_float(v[1] << 16), __uint_as_float(v[1] & 0xffff0000u)};
            a1 = fvec4{__uint_as_float(v[2] << 16), __uint_as_float(v[2] & 0xffff0000u),
                       __uint_as_float(v[3] << 16), __uint_as_float(v[3] & 0xffff0000u)};
        }
    };
    auto bstore = [&](const Stage &S, int b) {
#pragma unroll
        for (int u = 0; u < PP; u++) {
            const int p = tid + u * NT;
            if (p >= 256) break;
            const int kp = p / 16, m = p % 16;
            const int word = 4 * ((kp >> 2) ^ ((m >> 2) & 3)) + (kp & 3);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                uint32_t h, mm, l;
                bf_split3(S.b0[u][i], S.b1[u][i], h, mm, l);
                const int o = (4 * m + i) * kBfRW + word;
                bs[b][0][o] = h;
                bs[b][1][o] = mm;
                bs[b][2][o] = l;
            }
        }
    };
    v4d sum[4];
    fvec4 acc[4];
#pragma unroll
    for (int jb = 0; jb < 4; jb++) {
        sum[jb] = v4d{0.0, 0.0, 0.0, 0.0};
        acc[jb] = fvec4{0.f, 0.f, 0.f, 0.f};
    }
    // the six part products of one slice into the FP32 accumulators
    auto multiply = [&](int b, const Stage &S) {
        fvec4 a0, a1;
        widen_a(S, a0, a1);
        uint32_t ah[4], am[4], al[4];
        bf_split3(a0[0], a0[1], ah[0], am[0], al[0]);
        bf_split3(a0[2], a0[3], ah[1], am[1], al[1]);
        bf_split3(a1[0], a1[1], ah[2], am[2], al[2]);
        bf_split3(a1[2], a1[3], ah[3], am[3], al[3]);
        const bf16x8 Ah = as_bf8(uvec4{ah[0], ah[1], ah[2], ah[3]}), Am = as_bf8(uvec4{am[0], am[1], am[2], am[3]}),
                     Al = as_bf8(uvec4{al[0], al[1], al[2], al[3]});
        // two 16-column blocks of right-hand sides at a time: their B parts
        // (24 registers) are live, not all four blocks' (48)
#pragma unroll
        for (int jp = 0; jp < 4; jp += 2) {
            bf16x8 Bh[2], Bm[2], Bl[2];
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int jb = jp + u;
                const int o = (16 * jb + c) * kBfRW + 4 * (q ^ (jb & 3));
                Bh[u] = as_bf8(*(const uvec4 *)&bs[b][0][o]);
                Bm[u] = as_bf8(*(const uvec4 *)&bs[b][1][o]);
                Bl[u] = as_bf8(*(const uvec4 *)&bs[b][2][o]);
            }
#pragma unroll
            for (int u = 0; u < 2; u++) acc[jp + u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al, Bh[u], acc[jp + u], 0, 0, 0);
#pragma unroll
            for (int u = 0; u < 2; u++) acc[jp + u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bl[u], acc[jp + u], 0, 0, 0);
#pragma unroll
            for (int u = 0; u < 2; u++) acc[jp + u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Am, Bm[u], acc[jp + u], 0, 0, 0);
#pragma unroll
            for (int u = 0; u < 2; u++) acc[jp + u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Am, Bh[u], acc[jp + u], 0, 0, 0);
#pragma unroll
            for (int u = 0; u < 2; u++) acc[jp + u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bm[u], acc[jp + u], 0, 0, 0);
#pragma unroll
            for (int u = 0; u < 2; u++) acc[jp + u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bh[u], acc[jp + u], 0, 0, 0);
        }
    };
    // FP32 sums of at most two slices (64 columns) are added into the FP64 sums
    auto flush = [&] {
#pragma unroll
        for (int jb = 0; jb < 4; jb++) {
#pragma unroll
            for (int i = 0; i < 4; i++) sum[jb][i] += (double)acc[jb][i];
            acc[jb] = fvec4{0.f, 0.f, 0.f, 0.f};
        }
    };
#pragma unroll
    for (int b = 0; b < NS - 1; b++) load(st[b], b);
    bstore(st[0], 0);
    __syncthreads();
    // slice sl: stage sl % NS (i is a multiple of NS), LDS buffer sl % NB
    for (int i = 0; i < nsl; i += NS) {
#pragma unroll
        for (int b = 0; b < NS; b++) {
            const int sl = i + b;
            load(st[(b + NS - 1) % NS], sl + NS - 1);
            if (sl < nsl && live) multiply(NB == 2 ? (sl & 1) : 0, st[b]);
            if constexpr (NB == 1) __syncthreads();   // one buffer: every wave has read it
            bstore(st[(b + 1) % NS], NB == 2 ? ((sl + 1) & 1) : 0);
            __syncthreads();
            if (sl & 1) flush();
        }
    }
    flush();
    // rows 16 w + 4 q + i, right-hand sides 16 jb + c
    if constexpr (P2) {
        double *o = out + (size_t)it2.part * (16 * NW) * kRR;
#pragma unroll
        for (int jb = 0; jb < 4; jb++)
#pragma unroll
            for (int i = 0; i < 4; i++) o[(size_t)(16 * w + 4 * q + i) * kRR + 16 * jb + c] = sum[jb][i];
    } else {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int row = 16 * w + 4 * q + i;
            if (row >= it1.kt) continue;
#pragma unroll
            for (int jb = 0; jb < 4; jb++) {
                if (it1.part < 0) Zout[(size_t)zrow[it1.zr + row] * kRR + 16 * jb + c] = (float)sum[jb][i];
                else out[((size_t)it1.part * prow + row) * kRR + 16 * jb + c] = sum[jb][i];
            }
        }
    }
}

// Lane-order copy of phase-2 factors: block b of the table is (item, slice,
// dtype); its 16 NW rows x 32 columns are written as 64-lane x 8 groups per
// 16-row block (zeros past the item's columns)
struct MPack {
    uint64_t src, dst;   // byte offset in the plan buffer, float index in the copy
    int ncol, s, dtype, nw;
};
__global__ __launch_bounds__(256) void mr_pack_bf_kernel(const MPack *__restrict__ tab, const uint8_t *__restrict__ buf,
                                                         float *__restrict__ dst)
{
    const MPack t = tab[blockIdx.x];
    const int tr = 16 * t.nw;
    for (int e = threadIdx.x; e < t.nw * 64 * 8; e += blockDim.x) {
        const int rb = e / 512, lane = (e / 8) % 64, j = e % 8;
        const int row = 16 * rb + lane % 16, col = 32 * t.s + 8 * (lane / 16) + j;
        float v = 0.f;
        if (col < t.ncol) {
            const size_t k = (size_t)col * tr + row;
            v = t.dtype == kF32 ? ((const float *)(buf + t.src))[k]
                                : __uint_as_float((uint32_t)((const uint16_t *)(buf + t.src))[k] << 16);
        }
        dst[t.dst + ((size_t)t.s * t.nw * 64) * 8 + e] = v;
    }
}

// ------------------------------------------------------------ host side --
struct MrPlan {
    bool built = false;
    int kc1 = 512, cols2 = 256;    // phase-1 column chunk, phase-2 columns per item (autotuned)
    bool xcd1 = false;             // BF16 phase 1: items of one X chunk on one XCD (autotuned)
    int p1ns = kBfStages;          // BF16 phase 1: register stages of the load ring (autotuned)
    int p2ns = kBfStages;          // BF16 phase 2 (64-row tiles): the same (autotuned)
    bool f32 = false;              // FP32 / BF16 factors only: the FP32-core kernels, X32 and Z in FP32
    int p1rows = kP1Rows;          // a1^T rows per phase-1 item (kP1RowsBf on the BF16 cores)
    int r2[3][2] = {};             // phase-2 item range per dtype
    std::vector<MItem1> it1;
    std::vector<MBlk1> blk1;
    std::vector<Item2> it2;
    std::vector<MItem2> it2f;      // the same items as contiguous column blocks (FP32-core kernel)
    std::vector<int> cmap;         // their columns' X / Z rows (bit 31: Z)
    std::vector<int> part0;        // ntiles + 1: partial slots of each row tile
    DevBuf<MItem1> d_it1;
    DevBuf<MBlk1> d_blk1;
    std::vector<int2> zq;          // (block of blk1, first of 4 rows) per workgroup of the Z reduction
    DevBuf<int2> d_zq;
    DevBuf<Item2> d_it2;
    DevBuf<MItem2> d_it2f;
    DevBuf<int> d_cmap;
    DevBuf<int> d_part0;
    int nzpart = 0, nparts = 0;
    // BF16 matrix cores (FP32 / BF16 factors, split operands)
    bool bf = false;
    int n1f = 0;                   // BF16 cores: phase-1 items [0, n1f) have FP32 rows, the rest BF16 rows
    std::vector<MItem2b> it2b;     // phase-2 items with their lane-order factor copies
    int ncol2max = 0;              // columns of the widest phase-2 item (its column map is staged in LDS)
    std::vector<MPack> pack;       // (item, slice) blocks of the copy
    size_t a16_floats = 0;
    DevBuf<MItem2b> d_it2b;
    DevBuf<MPack> d_pack;
    DevBuf<float> a16;
};

}  // namespace hbi

namespace {
// per (matrix, precision) multi-RHS item lists, kept beside the matrix
struct MrCache {
    const hbi_hmat *h;
    int prec;
    hbi::MrPlan P;
};
std::vector<MrCache *> g_mr;
}  // namespace

namespace hbi {

// Workgroup b of a launch runs on XCD b % 8 (round-robin dispatch) and each
// XCD has its own L2.  Phase-1 items that read the same X rows (one column
// chunk of one column cluster: the row groups of its stacked a1^T rows) are
// placed on one XCD at consecutive positions, so the chunk is fetched into
// that L2 once; the chunks are dealt largest first to the XCD with the least
// MFMA work so far.  Positions left at the end of the shorter XCD lists hold
// empty items (no rows, no columns).
static void xcd_order_items1(std::vector<MItem1> &items)
{
    if (items.empty()) return;
    constexpr int kX = 8;
    auto work = [](const MItem1 &x) { return (long long)((x.kt + 15) / 16) * ((x.tn + 31) / 32); };
    std::map<std::pair<int, int>, int> gid;
    std::vector<std::vector<MItem1>> grp;
    std::vector<long long> gw;
    for (const MItem1 &it : items) {
        const auto key = std::make_pair(it.xoff + it.t0, it.tn);
        auto f = gid.find(key);
        int g;
        if (f == gid.end()) {
            g = (int)grp.size();
            gid.emplace(key, g);
            grp.emplace_back();
            gw.push_back(0);
        } else {
            g = f->second;
        }
        grp[g].push_back(it);
        gw[g] += work(it);
    }
    std::vector<int> ord(grp.size());
    for (size_t g = 0; g < ord.size(); g++) ord[g] = (int)g;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return gw[a] > gw[b]; });
    std::vector<MItem1> lane[kX];
    long long lw[kX] = {};
    for (int g : ord) {
        int x = 0;
        for (int k = 1; k < kX; k++)
            if (lw[k] < lw[x]) x = k;
        lane[x].insert(lane[x].end(), grp[g].begin(), grp[g].end());
        lw[x] += gw[g];
    }
    size_t len = 0;
    for (int x = 0; x < kX; x++) len = std::max(len, lane[x].size());
    MItem1 empty{};
    empty.part = -1;
    items.assign(len * kX, empty);
    for (int x = 0; x < kX; x++)
        for (size_t k = 0; k < lane[x].size(); k++) items[k * kX + x] = lane[x][k];
}

// Phase-1 items.  Runs of blocks with the same column cluster and dtype are
// adjacent in the packed layout (same ldt): whole blocks are packed into row
// groups of at most M.p1rows rows, each read against its x rows once per
// column chunk of kc1.  Ranks above it are split into pieces of their own.
static void mr_items1(const Plan &P, MrPlan &M, int kc1)
{
    const int cap = M.p1rows;
    M.it1.clear();
    M.blk1.clear();
    M.nzpart = 0;
    const auto &R = P.rows1;
    // stacked row of every rows1 entry (rows1 and the z map follow one order)
    std::vector<int> srow(R.size() + 1, 0);
    for (size_t q = 0; q < R.size(); q++) srow[q + 1] = srow[q] + R[q].kt;
    size_t i = 0;
    while (i < R.size()) {
        struct Piece { uint64_t off; int kt, zoff, sr; };
        std::vector<Piece> grp;
        int rows = 0;
        const Row1 &f = R[i];
        if (f.kt > cap) {
            for (int k0 = 0; k0 < f.kt; k0 += cap)
                grp.push_back({f.off + (uint64_t)k0 * f.ldt * dtype_size(f.dtype), std::min(cap, f.kt - k0),
                               f.zoff + k0, srow[i] + k0});
            i++;
        } else {
            size_t j = i;
            while (j < R.size() && R[j].xoff == f.xoff && R[j].nc == f.nc && R[j].dtype == f.dtype &&
                   R[j].kt <= cap && rows + R[j].kt <= cap &&
                   (j == i || R[j].off == R[j - 1].off + (uint64_t)R[j - 1].kt * R[j - 1].ldt * dtype_size(f.dtype))) {
                grp.push_back({R[j].off, R[j].kt, R[j].zoff, srow[j]});
                rows += R[j].kt;
                j++;
            }
            i = j;
        }
        // one packed run per group, except split pieces: each piece is its own run
        const bool split = f.kt > cap;
        for (size_t g0 = 0; g0 < grp.size();) {
            const size_t g1 = split ? g0 + 1 : grp.size();
            int kt = 0;
            for (size_t g = g0; g < g1; g++) kt += grp[g].kt;
            const bool direct = f.nc <= kc1;
            const int p0 = M.nzpart;
            for (int t0 = 0; t0 < f.nc; t0 += kc1) {
                MItem1 it{};
                it.off = grp[g0].off; it.ldt = f.ldt; it.kt = kt; it.t0 = t0; it.tn = std::min(kc1, f.nc - t0);
                it.xoff = f.xoff; it.part = direct ? -1 : M.nzpart++; it.dtype = f.dtype; it.zr = grp[g0].sr;
                it.r0 = 0;
                M.it1.push_back(it);
            }
            if (!direct) {
                int roff = 0;
                for (size_t g = g0; g < g1; g++) {
                    M.blk1.push_back({p0, M.nzpart, grp[g].kt, grp[g].zoff, roff});
                    roff += grp[g].kt;
                }
            }
            g0 = g1;
        }
    }
    // launch order: most MFMA work first; partial slots keep their numbers
    std::stable_sort(M.it1.begin(), M.it1.end(), [](const MItem1 &x, const MItem1 &y) {
        return (x.kt + 15) / 16 * ((x.tn + 31) / 32) > (y.kt + 15) / 16 * ((y.tn + 31) / 32);
    });
    if (M.bf) {
        // BF16 cores: every item fits 4 waves (kP1RowsBf rows); the items of
        // FP32 rows and those of BF16 rows are two launches (a kernel for
        // each row type uses fewer registers), each in its own XCD order
        std::vector<MItem1> f32, b16;
        for (const MItem1 &it : M.it1) (it.dtype == kF32 ? f32 : b16).push_back(it);
        if (M.xcd1) {
            xcd_order_items1(f32);
            xcd_order_items1(b16);
        }
        M.n1f = (int)f32.size();
        M.it1 = f32;
        M.it1.insert(M.it1.end(), b16.begin(), b16.end());
    }
    M.kc1 = kc1;
}

// Phase-2 items: runs of one row tile's segments of one dtype with about
// cols2 columns in total (the single-vector items are sized in bytes for an
// HBM-bound kernel; these are sized for MFMA work per workgroup).
static void mr_items2(const hbi_hmat *h, const Plan &P, MrPlan &M, int cols2)
{
    M.it2.clear();
    M.part0.assign(h->ntiles + 1, 0);
    int nparts = 0;
    size_t g = 0;
    for (int t = 0; t < h->ntiles; t++) {
        M.part0[t] = nparts;
        for (; g < P.groups.size() && P.groups[g].tile == t; g++) {
            const SegGroup &G = P.groups[g];
            int cur = 0, s0 = G.s0;
            for (int q = G.s0; q < G.s1; q++) {
                if (q > s0 && cur + P.segs[q].ncol > cols2) {
                    M.it2.push_back({t, s0, q, nparts++, G.dtype, 0});
                    s0 = q;
                    cur = 0;
                }
                cur += P.segs[q].ncol;
            }
            M.it2.push_back({t, s0, G.s1, nparts++, G.dtype, 0});
        }
    }
    M.part0[h->ntiles] = nparts;
    M.nparts = nparts;
    std::stable_sort(M.it2.begin(), M.it2.end(), [](const Item2 &x, const Item2 &y) { return x.dtype < y.dtype; });
    for (int d = 0; d < 3; d++) M.r2[d][0] = M.r2[d][1] = 0;
    for (int q = 0; q < (int)M.it2.size(); q++) {
        const int d = M.it2[q].dtype;
        if (M.r2[d][1] == 0) M.r2[d][0] = q;
        M.r2[d][1] = q + 1;
    }
    // contiguous-block form: the segments of an item follow each other in the layout
    M.it2f.clear();
    M.cmap.clear();
    for (const Item2 &it : M.it2) {
        MItem2 f{};
        f.off = P.segs[it.seg0].off;
        f.cm0 = (int)M.cmap.size();
        f.part = it.part;
        const size_t colbytes = (size_t)h->tr * dtype_size(it.dtype);
        for (int q = it.seg0; q < it.seg1; q++) {
            const Seg &g = P.segs[q];
            HBI_REQUIRE(g.off == f.off + (uint64_t)f.ncol * colbytes, "64-column product: item segments not contiguous");
            for (int j = 0; j < g.ncol; j++) M.cmap.push_back((g.src ? (int)0x80000000u : 0) | (g.xoff + j));
            f.ncol += g.ncol;
        }
        HBI_REQUIRE(f.ncol <= kMaxCols32, "64-column product: item wider than the column map");
        M.it2f.push_back(f);
    }
    if (M.bf) {
        // lane-order copies: 16 x 32 blocks of 64 lanes x 8 floats
        const int nw = h->tr / 16;
        M.it2b.clear();
        M.ncol2max = 0;
        M.pack.clear();
        uint64_t off = 0;
        for (size_t q = 0; q < M.it2.size(); q++) {
            const MItem2 &f = M.it2f[q];
            M.it2b.push_back({off, f.ncol, f.cm0, f.part, 0});
            M.ncol2max = std::max(M.ncol2max, f.ncol);
            const int nsl = (f.ncol + 31) / 32;
            for (int sl = 0; sl < nsl; sl++) M.pack.push_back({f.off, off, f.ncol, sl, M.it2[q].dtype, nw});
            off += (uint64_t)nsl * nw * 64 * 8;
        }
        M.a16_floats = off;
    }
    M.cols2 = cols2;
}

// Core selection: a plan holds either FP64 factors only or none (the
// precision policies never mix them), so Z has one type.  FP32 / BF16
// factors of HBI_PREC_MIXED multiply on the BF16 cores (split operands, the
// faster form), those of HBI_PREC_F32 on the FP32 cores.  The two differ in
// their rounding: the FP32 cores round each addition to nearest, while the
// BF16 cores align the products of each group of 8 columns to the largest and
// truncate the parts of the others below half an FP32 unit of it (toward
// zero), a biased error of the same size that ensembles integrated over many
// steps accumulate (event onsets ~1e-5 relative from FP64 against ~1e-6).
static void mr_mode(const Plan &P, MrPlan &M, int prec)
{
    bool has64 = false, has32 = false;
    for (const Seg &g : P.segs) (g.dtype == kF64 ? has64 : has32) = true;
    for (const Row1 &r : P.rows1) (r.dtype == kF64 ? has64 : has32) = true;
    HBI_REQUIRE(!(has64 && has32), "64-column product: FP64 and reduced-precision factors in one plan");
    for (const Seg &g : P.segs)
        HBI_REQUIRE(g.off % 16 == 0, "64-column product: segment not 16-byte aligned");
    M.f32 = has32;
    M.bf = has32 && prec == HBI_PREC_MIXED;
    M.p1rows = M.bf ? kP1RowsBf : (has32 ? kP1Rows32 : kP1Rows);
}

static void mr_upload(hbi_hmat *h, const Plan &P, MrPlan &M)
{
    M.d_it1.upload(M.it1, h->stream);
    M.d_blk1.upload(M.blk1, h->stream);
    M.zq.clear();
    for (size_t b = 0; b < M.blk1.size(); b++)
        for (int r = 0; r < M.blk1[b].kt; r += 4) M.zq.push_back(int2{(int)b, r});
    M.d_zq.upload(M.zq, h->stream);
    M.d_it2.upload(M.it2, h->stream);
    M.d_it2f.upload(M.it2f, h->stream);
    M.d_cmap.upload(M.cmap.empty() ? std::vector<int>{0} : M.cmap, h->stream);
    M.d_part0.upload(M.part0, h->stream);
    if (M.bf && !M.pack.empty()) {
        M.d_it2b.upload(M.it2b, h->stream);
        M.d_pack.upload(M.pack, h->stream);
        if (M.a16.n < M.a16_floats) M.a16.alloc(M.a16_floats);
        mr_pack_bf_kernel<<<(int)M.pack.size(), 256, 0, h->stream>>>(M.d_pack.p, P.buf.p, M.a16.p);
        HBI_HIP(hipGetLastError());
        HBI_HIP(hipStreamSynchronize(h->stream));
    }
}

static void mr_tune(hbi_hmat *h, int prec, MrPlan &M);

const MrPlan &mr_plan(hbi_hmat *h, int prec)
{
    for (MrCache *c : g_mr)
        if (c->h == h && c->prec == prec) return c->P;
    const Plan &P = hmat_plan(h, prec);
    MrCache *c = new MrCache{h, prec, MrPlan()};
    try {
        MrPlan &M = c->P;
        mr_mode(P, M, prec);
        mr_items1(P, M, 512);
        mr_items2(h, P, M, 256);
        mr_upload(h, P, M);
        M.built = true;
        mr_tune(h, prec, M);
    } catch (...) {
        delete c;
        throw;
    }
    g_mr.push_back(c);
    return c->P;
}

void mr_forget(const hbi_hmat *h)
{
    for (size_t i = 0; i < g_mr.size();) {
        if (g_mr[i]->h == h) {
            delete g_mr[i];
            g_mr.erase(g_mr.begin() + i);
        } else {
            i++;
        }
    }
}

// buffer sizes cover every candidate of the tuner
static size_t zpart_cap(const MrPlan &M) { return (size_t)std::max(M.nzpart, 1) * M.p1rows * kRR; }
static size_t ypart_cap(const hbi_hmat *h, const MrPlan &M) { return (size_t)std::max(M.nparts, 1) * h->tr * kRR; }
size_t mr_zpart_size(hbi_hmat *h, int prec) { return zpart_cap(mr_plan(h, prec)); }
size_t mr_x32_size(hbi_hmat *h, int prec) { return mr_plan(h, prec).f32 ? (size_t)h->n * kRR : 1; }
size_t mr_ypart_size(hbi_hmat *h, int prec) { return ypart_cap(h, mr_plan(h, prec)); }
const int *mr_part0(hbi_hmat *h, int prec) { return mr_plan(h, prec).d_part0.p; }

// Z = a1^T X for every low-rank block (partials, then the ordered reduction);
// in FP32-core mode X is first rounded into X32 (which phase 2 reads too) and
// Z is stored as FP32 in the first half of its buffer
static void phase1(hbi_hmat *h, const Plan &P, const MrPlan &M, const double *X, float *X32, double *zpart,
                   double *Z, hipStream_t s, const int *gate)
{
    if (M.f32 && X) {   // X null: the caller wrote X32 itself
        const size_t n4 = (size_t)h->n * kRR / 4;
        mr_narrow_kernel<<<(int)((n4 + 255) / 256), 256, 0, s>>>(n4, X, X32, gate);
    }
    if (M.it1.empty()) return;
    if (!M.f32) {
        mr_lrvt_kernel<<<(int)M.it1.size(), 64, 0, s>>>(M.d_it1.p, P.buf.p, X, zpart, P.d_zrow.p, Z, gate);
        if (!M.zq.empty())
            mr_zreduce_kernel<double><<<(int)M.zq.size(), 4 * kRR, 0, s>>>(M.d_blk1.p, M.d_zq.p,
                                                                                   M.p1rows, zpart, Z, gate);
    } else if (M.bf) {
        float *Z32 = (float *)Z;
        // every item fits 4 waves (kP1RowsBf rows); 2 or 3 register stages (autotuned); items of FP32
        // rows, then those of BF16 rows
        const int nf = M.n1f, nb = (int)M.it1.size() - M.n1f;
        const MItem1 *i16 = M.d_it1.p + nf;
        if (M.p1ns == 3) {
            if (nf)
                mr_bf16_kernel<4, false, 3, kBfBufs><<<nf, 256, 0, s>>>(
                    nullptr, M.d_it1.p, nullptr, nullptr, P.buf.p, X32, nullptr, zpart, P.d_zrow.p, Z32, M.p1rows, gate);
            if (nb)
                mr_bf16_kernel<4, false, 3, kBfBufs, true><<<nb, 256, 0, s>>>(
                    nullptr, i16, nullptr, nullptr, P.buf.p, X32, nullptr, zpart, P.d_zrow.p, Z32, M.p1rows, gate);
        } else {
            if (nf)
                mr_bf16_kernel<4, false, kBfStages, kBfBufs><<<nf, 256, 0, s>>>(
                    nullptr, M.d_it1.p, nullptr, nullptr, P.buf.p, X32, nullptr, zpart, P.d_zrow.p, Z32, M.p1rows, gate);
            if (nb)
                mr_bf16_kernel<4, false, kBfStages, kBfBufs, true><<<nb, 256, 0, s>>>(
                    nullptr, i16, nullptr, nullptr, P.buf.p, X32, nullptr, zpart, P.d_zrow.p, Z32, M.p1rows, gate);
        }
        if (!M.zq.empty())
            mr_zreduce_kernel<float><<<(int)M.zq.size(), 4 * kRR, 0, s>>>(M.d_blk1.p, M.d_zq.p,
                                                                                  M.p1rows, zpart, Z32, gate);
    } else {
        float *Z32 = (float *)Z;
        mr_lrvt32_kernel<<<(int)M.it1.size(), 2 * kP1Rows32, 0, s>>>(M.d_it1.p, P.buf.p, X32, zpart, P.d_zrow.p, Z32,
                                                                     gate);
        if (!M.zq.empty())
            mr_zreduce_kernel<float><<<(int)M.zq.size(), 4 * kRR, 0, s>>>(M.d_blk1.p, M.d_zq.p,
                                                                                  M.p1rows, zpart, Z32, gate);
    }
}

// Y partials (per tile item) of the dense and a2 Z products
static void phase2(const hbi_hmat *h, const Plan &P, const MrPlan &M, const double *X, const float *X32,
                   const double *Z, double *ypart, hipStream_t s, const int *gate)
{
    const float *Z32 = (const float *)Z;
    if (M.bf) {   // every FP32 / BF16 item in one launch (the copies are FP32)
        const int n = (int)M.it2b.size();
        if (n == 0) return;
        const size_t cmb = (size_t)M.ncol2max * sizeof(int);
        if (h->tr == 64 && M.p2ns == 3)
            mr_bf16_kernel<4, true, 3, kBfBufs><<<n, 256, cmb, s>>>(M.d_it2b.p, nullptr, M.d_cmap.p, M.a16.p, nullptr,
                                                                   X32, Z32, ypart, nullptr, nullptr, 0, gate);
        else if (h->tr == 64)
            mr_bf16_kernel<4, true, kBfStages, kBfBufs><<<n, 256, cmb, s>>>(M.d_it2b.p, nullptr, M.d_cmap.p, M.a16.p, nullptr, X32,
                                                                 Z32, ypart, nullptr, nullptr, 0, gate);
        else
            mr_bf16_kernel<2, true, kBfStages, kBfBufs><<<n, 128, cmb, s>>>(M.d_it2b.p, nullptr, M.d_cmap.p, M.a16.p, nullptr, X32,
                                                                 Z32, ypart, nullptr, nullptr, 0, gate);
        return;
    }
    for (int d = 0; d < 3; d++) {
        const int n = M.r2[d][1] - M.r2[d][0];
        if (n <= 0) continue;
        const Item2 *it = M.d_it2.p + M.r2[d][0];
        const MItem2 *itf = M.d_it2f.p + M.r2[d][0];
        const int *cm = M.d_cmap.p;
        if (h->tr == 64) {
            if (d == kF64) mr_tile_kernel<double, 64><<<n, kMrThreads, 0, s>>>(it, P.d_segs.p, P.buf.p, X, Z, ypart, gate);
            else if (d == kF32) mr_tile32_kernel<float, 64, 4><<<n, 512, 0, s>>>(itf, cm, P.buf.p, X32, Z32, ypart, gate);
            else mr_tile32_kernel<uint16_t, 64, 4><<<n, 512, 0, s>>>(itf, cm, P.buf.p, X32, Z32, ypart, gate);
        } else {
            if (d == kF64) mr_tile_kernel<double, 32><<<n, kMrThreads, 0, s>>>(it, P.d_segs.p, P.buf.p, X, Z, ypart, gate);
            else if (d == kF32) mr_tile32_kernel<float, 32, 2><<<n, 256, 0, s>>>(itf, cm, P.buf.p, X32, Z32, ypart, gate);
            else mr_tile32_kernel<uint16_t, 32, 2><<<n, 256, 0, s>>>(itf, cm, P.buf.p, X32, Z32, ypart, gate);
        }
    }
}

void mr_phase1(hbi_hmat *h, int prec, const double *X, float *X32, double *zpart, double *Z, hipStream_t s,
               const int *gate)
{
    phase1(h, hmat_plan(h, prec), mr_plan(h, prec), X, X32, zpart, Z, s, gate);
}

void mr_phase2(hbi_hmat *h, int prec, const double *X, const float *X32, const double *Z, double *ypart,
               hipStream_t s, const int *gate)
{
    phase2(h, hmat_plan(h, prec), mr_plan(h, prec), X, X32, Z, ypart, s, gate);
}

// Y partials of H X (device pointers, row-major N x 64); Y itself through mr_reduce or a consumer.
// X32 (mr_x32_size floats) is scratch for the FP32 copy of X; with the FP32
// cores a caller may fill X32 itself and pass X = null.
void mr_apply_parts(hbi_hmat *h, int prec, const double *X, float *X32, double *zpart, double *Z, double *ypart,
                    hipStream_t s, const int *gate)
{
    mr_phase1(h, prec, X, X32, zpart, Z, s, gate);
    mr_phase2(h, prec, X, X32, Z, ypart, s, gate);
}

void mr_reduce(hbi_hmat *h, int prec, const double *ypart, double *Y, hipStream_t s)
{
    const MrPlan &M = mr_plan(h, prec);
    const size_t tot = (size_t)h->n * kRR;
    mr_tile_reduce_kernel<<<(int)((tot + 255) / 256), 256, 0, s>>>(h->n, h->tr, M.d_part0.p, ypart, Y);
}

}  // namespace hbi

namespace hbi {

// Pick the phase-1 column chunk and the phase-2 columns per item by timing
// the product on this matrix (zero input; the kernels' work does not depend
// on the values).
static void mr_tune(hbi_hmat *h, int prec, MrPlan &M)
{
    const Plan &P = hmat_plan(h, prec);
    const int kc1s[] = {64, 128, 256, 512, 1024, 2048};
    const int c2s[] = {128, 256, 512, 1024, 2048, 4096};
    // buffers sized for the smallest chunks (most partial slots)
    mr_items1(P, M, kc1s[0]);
    mr_items2(h, P, M, c2s[0]);
    DevBuf<double> dX((size_t)h->n * kRR), dZ((size_t)std::max(h->nz, 1) * kRR);
    DevBuf<double> zp(zpart_cap(M)), yp(ypart_cap(h, M)), dY((size_t)h->n * kRR);
    DevBuf<float> x32(M.f32 ? (size_t)h->n * kRR : 1);
    HBI_HIP(hipMemsetAsync(dX.p, 0, dX.n * 8, h->stream));
    HBI_HIP(hipMemsetAsync(dZ.p, 0, dZ.n * 8, h->stream));
    auto time_it = [&](int which) {
        mr_upload(h, P, M);
        // phase 2 is timed with the reduction of its partial slots, which every
        // consumer pays (fewer, larger items: fewer slots to read)
        auto run = [&] {
            if (which == 1) {
                phase1(h, P, M, dX.p, x32.p, zp.p, dZ.p, h->stream, nullptr);
            } else {
                phase2(h, P, M, dX.p, x32.p, dZ.p, yp.p, h->stream, nullptr);
                const size_t tot = (size_t)h->n * kRR;
                mr_tile_reduce_kernel<<<(int)((tot + 255) / 256), 256, 0, h->stream>>>(h->n, h->tr, M.d_part0.p,
                                                                                        yp.p, dY.p);
            }
        };
        for (int w = 0; w < 2; w++) run();
        HBI_HIP(hipEventRecord(h->ev[0], h->stream));
        for (int r = 0; r < 5; r++) run();
        HBI_HIP(hipEventRecord(h->ev[1], h->stream));
        HBI_HIP(hipEventSynchronize(h->ev[1]));
        float ms = 0;
        HBI_HIP(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
        return ms;
    };
    int b1 = 512, b2 = 256, bn = kBfStages;
    float best = 1e30f;
    for (int ns : {kBfStages, 3}) {
        if (ns == 3 && !(M.bf && h->tr == 64)) break;
        M.p2ns = ns;
        for (int c : c2s) {
            mr_items2(h, P, M, c);
            const float ms = time_it(2);
            if (ms < best) { best = ms; b2 = c; bn = ns; }
        }
    }
    M.p2ns = bn;
    // phase 1: column chunk for each item order; the XCD placement fetches
    // about half the bytes from beyond L2 (the X chunks once per XCD) and is
    // kept unless it is more than 2% slower
    float bt[2] = {1e30f, 1e30f};
    int bc[2] = {512, 512};
    for (int x = 0; x < (M.bf ? 2 : 1); x++)
        for (int c : kc1s) {
            M.xcd1 = x == 1;
            mr_items1(P, M, c);
            const float ms = time_it(1);
            if (ms < bt[x]) { bt[x] = ms; bc[x] = c; }
        }
    M.xcd1 = M.bf && bt[1] <= 1.02f * bt[0];
    b1 = bc[M.xcd1 ? 1 : 0];
    mr_items1(P, M, b1);
    if (M.bf) {   // a deeper load ring for phase 1 at the chosen items
        const float t2 = time_it(1);
        M.p1ns = 3;
        if (time_it(1) >= t2) M.p1ns = kBfStages;
    }
    mr_items2(h, P, M, b2);
    mr_upload(h, P, M);
}

}  // namespace hbi

extern "C" int hbi_hmat_matmul(hbi_hmat *h, const double *X, double *Y, int nrhs, int prec)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(h && X && Y && nrhs > 0, "bad arguments");
        const int n = h->n;
        const hbi::Plan &P = hbi::hmat_plan(h, prec);
        (void)P;
        hbi::DevBuf<double> dX((size_t)n * hbi::kRR), dY((size_t)n * hbi::kRR), dZ((size_t)(std::max(h->nz, 1)) * hbi::kRR);
        hbi::DevBuf<double> zp(hbi::mr_zpart_size(h, prec)), yp(hbi::mr_ypart_size(h, prec));
        hbi::DevBuf<float> x32(hbi::mr_x32_size(h, prec));
        std::vector<double> hx((size_t)n * hbi::kRR), hy((size_t)n * hbi::kRR);
        for (int c0 = 0; c0 < nrhs; c0 += hbi::kRR) {
            const int nc = std::min(hbi::kRR, nrhs - c0);
            std::fill(hx.begin(), hx.end(), 0.0);
            for (int r = 0; r < nc; r++)
                for (int i = 0; i < n; i++) hx[(size_t)i * hbi::kRR + r] = X[(size_t)(c0 + r) * n + i];
            dX.upload(hx, h->stream);
            hbi::mr_apply_parts(h, prec, dX.p, x32.p, zp.p, dZ.p, yp.p, h->stream, nullptr);
            hbi::mr_reduce(h, prec, yp.p, dY.p, h->stream);
            HBI_HIP(hipGetLastError());
            dY.download(hy.data(), hy.size(), h->stream);
            HBI_HIP(hipStreamSynchronize(h->stream));
            for (int r = 0; r < nc; r++)
                for (int i = 0; i < n; i++) Y[(size_t)(c0 + r) * n + i] = hy[(size_t)i * hbi::kRR + r];
        }
    });
}

extern "C" int hbi_hmat_matmul_plan(hbi_hmat *h, int prec, int *info)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(h && info, "bad arguments");
        const hbi::MrPlan &M = hbi::mr_plan(h, prec);
        info[0] = M.kc1;
        info[1] = M.cols2;
        info[2] = (int)M.it1.size();
        info[3] = (int)M.it2.size();
        info[4] = M.nzpart;
        info[5] = M.bf ? 2 : (M.f32 ? 1 : 0);
        info[6] = (M.xcd1 ? 1 : 0) | (M.p1ns << 1) | (M.p2ns << 4);
    });
}

extern "C" int hbi_hmat_time_matmul(hbi_hmat *h, int prec, int iters, double *ms)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(h && ms && iters > 0, "bad arguments");
        const int n = h->n;
        const hbi::Plan &P = hbi::hmat_plan(h, prec);
        hbi::DevBuf<double> dX((size_t)n * hbi::kRR), dY((size_t)n * hbi::kRR), dZ((size_t)(std::max(h->nz, 1)) * hbi::kRR);
        hbi::DevBuf<double> zp(hbi::mr_zpart_size(h, prec)), yp(hbi::mr_ypart_size(h, prec));
        hbi::DevBuf<float> x32(hbi::mr_x32_size(h, prec));
        HBI_HIP(hipMemsetAsync(dX.p, 0, dX.n * 8, h->stream));
        HBI_HIP(hipMemsetAsync(dZ.p, 0, dZ.n * 8, h->stream));
        for (int w = 0; w < 2; w++) hbi::mr_apply_parts(h, prec, dX.p, x32.p, zp.p, dZ.p, yp.p, h->stream, nullptr);
        // the phases as they run inside whole products (phase 2 after phase 1,
        // so neither finds its own data of the previous call in the caches):
        // events between them, summed over the calls
        std::vector<hipEvent_t> ev(3 * (size_t)iters, nullptr);
        struct Events {
            std::vector<hipEvent_t> &v;
            ~Events()
            {
                for (hipEvent_t e : v)
                    if (e) (void)hipEventDestroy(e);
            }
        } guard{ev};
        for (auto &e : ev) HBI_HIP(hipEventCreate(&e));
        for (int it = 0; it < iters; it++) {
            HBI_HIP(hipEventRecord(ev[3 * it], h->stream));
            hbi::mr_phase1(h, prec, dX.p, x32.p, zp.p, dZ.p, h->stream, nullptr);
            HBI_HIP(hipEventRecord(ev[3 * it + 1], h->stream));
            hbi::mr_phase2(h, prec, dX.p, x32.p, dZ.p, yp.p, h->stream, nullptr);
            HBI_HIP(hipEventRecord(ev[3 * it + 2], h->stream));
            hbi::mr_reduce(h, prec, yp.p, dY.p, h->stream);
        }
        HBI_HIP(hipStreamSynchronize(h->stream));
        double t1 = 0.0, t2 = 0.0;
        for (int it = 0; it < iters; it++) {
            float a = 0, b = 0;
            HBI_HIP(hipEventElapsedTime(&a, ev[3 * it], ev[3 * it + 1]));
            HBI_HIP(hipEventElapsedTime(&b, ev[3 * it + 1], ev[3 * it + 2]));
            t1 += a;
            t2 += b;
        }
        ms[0] = t1 / iters;
        ms[1] = t2 / iters;
        // whole products back to back
        HBI_HIP(hipEventRecord(h->ev[0], h->stream));
        for (int it = 0; it < iters; it++) {
            hbi::mr_apply_parts(h, prec, dX.p, x32.p, zp.p, dZ.p, yp.p, h->stream, nullptr);
            hbi::mr_reduce(h, prec, yp.p, dY.p, h->stream);
        }
        HBI_HIP(hipEventRecord(h->ev[1], h->stream));
        HBI_HIP(hipEventSynchronize(h->ev[1]));
        float t = 0;
        HBI_HIP(hipEventElapsedTime(&t, h->ev[0], h->ev[1]));
        ms[2] = (double)t / iters;
        (void)P;
    });
}

// Rate of the matrix cores on this device for the instructions of the
// 64-column kernels: every wave runs iters rounds of 8 independent
// v_mfma_f64_16x16x4f64 (FP64 factors) or v_mfma_f32_16x16x4f32 (FP32 / BF16
// factors) chains, no memory traffic, waves_per_simd waves on each SIMD.  The
// ceiling the products are compared with beside the data-sheet peak.
namespace {
__global__ __launch_bounds__(64) void mfma_rate_kernel(double *out, int iters, double a0)
{
    using hbi::v4d;
    v4d acc[8];
#pragma unroll
    for (int i = 0; i < 8; i++) acc[i] = v4d{0.0, 0.0, 0.0, 0.0};
    const double a = a0 + threadIdx.x, b = 1.0 - a0 * threadIdx.x;
    for (int it = 0; it < iters; it++)
#pragma unroll
        for (int i = 0; i < 8; i++) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 8; i++) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
    out[(size_t)blockIdx.x * 64 + threadIdx.x] = s;
}

__global__ __launch_bounds__(64) void mfma32_rate_kernel(double *out, int iters, double a0)
{
    using hbi::fvec4;
    fvec4 acc[8];
#pragma unroll
    for (int i = 0; i < 8; i++) acc[i] = fvec4{0.f, 0.f, 0.f, 0.f};
    const float a = (float)(a0 + threadIdx.x), b = (float)(1.0 - a0 * threadIdx.x);
    for (int it = 0; it < iters; it++)
#pragma unroll
        for (int i = 0; i < 8; i++) acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[i], 0, 0, 0);
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 8; i++) s += (double)(acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3]);
    out[(size_t)blockIdx.x * 64 + threadIdx.x] = s;
}

int mfma_rate(bool f32, int waves_per_simd, int iters, double *tflops)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(tflops && waves_per_simd > 0 && waves_per_simd <= 8 && iters > 0, "bad arguments");
        int dev = 0, ncu = 0;
        HBI_HIP(hipGetDevice(&dev));
        HBI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        const int nwaves = ncu * 4 * waves_per_simd;
        hbi::DevBuf<double> out((size_t)nwaves * 64);
        hipStream_t s = nullptr;
        HBI_HIP(hipStreamCreate(&s));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HBI_HIP(hipEventCreate(&e0));
        HBI_HIP(hipEventCreate(&e1));
        auto launch = [&] {
            if (f32) mfma32_rate_kernel<<<nwaves, 64, 0, s>>>(out.p, iters, 0.5);
            else mfma_rate_kernel<<<nwaves, 64, 0, s>>>(out.p, iters, 0.5);
        };
        launch();
        HBI_HIP(hipEventRecord(e0, s));
        launch();
        HBI_HIP(hipEventRecord(e1, s));
        HBI_HIP(hipEventSynchronize(e1));
        float ms = 0;
        HBI_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
        *tflops = (double)nwaves * iters * 8 * 2048.0 / (ms * 1e-3) / 1e12;
    });
}
}  // namespace

extern "C" int hbi_mfma_f64_rate(int waves_per_simd, int iters, double *tflops)
{
    return mfma_rate(false, waves_per_simd, iters, tflops);
}

extern "C" int hbi_mfma_f32_rate(int waves_per_simd, int iters, double *tflops)
{
    return mfma_rate(true, waves_per_simd, iters, tflops);
}
