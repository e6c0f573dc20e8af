// multirhs.hip -- H-matrix times a block of RR = 64 right-hand sides.
//
// Ensembles of fault realisations share one H-matrix, so
// the per-stage product becomes Y (N x RR) = H X.  The factors are read once
// for all right-hand sides, which moves the product from HBM-bound to
// arithmetic-bound.  X, Z, Y are row-major with RR values per row (RR = 64,
// unused columns zero).
//
//   phase 1  Zpart(item) = a1^T[:, t-range] X[t-range, :] for column chunks of
//            every low-rank block (FP64 matrix cores); zreduce adds the
//            chunks of a block in order -> Z (sum_kt x RR)
//   phase 2  per tile item: Ypart = sum over segments of A_seg (TR x ncol)
//            X_seg (ncol x RR) on the FP64 matrix cores, 16-deep k-slices
//            staged through LDS
// Both phases reuse the packed single-vector layout of the precision's plan.
#include "hmat.hpp"

#include <algorithm>

namespace hbi {

constexpr int kRR = 64;      // right-hand sides per product
constexpr int kKS = 32;      // k-slice staged in LDS
constexpr int kMrThreads = 256;

template <typename T> __device__ __forceinline__ double widen(T v);
template <> __device__ __forceinline__ double widen<double>(double v) { return v; }
template <> __device__ __forceinline__ double widen<float>(float v) { return (double)v; }
template <> __device__ __forceinline__ double widen<uint16_t>(uint16_t v) { return (double)__uint_as_float(((uint32_t)v) << 16); }

// ------------------------------------------------------------- phase 1 --
struct MItem1 {
    uint64_t off;      // byte offset of the first a1^T row
    int ldt, kt;       // row stride (elements), rows (<= 64: whole blocks of one column cluster, stacked)
    int t0, tn;        // column chunk [t0, t0+tn)
    int xoff;          // first x row of the column cluster
    int part;          // partial slot: zpart[part][64][RR]
    int dtype;
};
struct MBlk1 {         // rows roff .. roff+kt of partial slots [p0, p1) -> z rows zoff .. zoff+kt
    int p0, p1, kt, zoff, roff;
};

typedef double v4d __attribute__((ext_vector_type(4)));

// One workgroup (4 waves) per item: Zpart (kt x 64) = a1^T[rows, t0..t0+tn]
// X[t0..t0+tn, :] on the FP64 matrix cores (16x16x4 MFMA, lane l: A[l % 16]
// [l / 16], B[l / 16][l % 16], result q of lane l at D[l / 16 + 4 q][l % 16]).
// Wave w owns right-hand sides [16 w, 16 w + 16) and the NRB = ceil(kt / 16)
// row blocks (items are launched grouped by NRB).  KS-deep slices of a1^T (transposed into LDS) and X pass through
// registers: the next slice's global loads are in flight while the current
// slice is multiplied.
template <typename T, int NRB>
__global__ __launch_bounds__(kMrThreads) void mr_lrvt_kernel(const MItem1 *__restrict__ items,
                                                             const uint8_t *__restrict__ buf,
                                                             const double *__restrict__ X, double *__restrict__ zpart,
                                                             const int *__restrict__ gate)
{
    constexpr int KS = 32;
    constexpr int AP = 64 + 1, XP = kRR + 16;   // odd AP: the transposing stores hit 32 distinct banks
    constexpr int NA = 64 * KS / kMrThreads, NX = KS * kRR / kMrThreads;
    __shared__ double as[KS][AP];     // as[t][k] = a1^T[k][t]
    __shared__ double xs[KS][XP];
    if (gate && *gate) return;
    const MItem1 it = items[blockIdx.x];
    const T *a = (const T *)(buf + it.off);
    const int tid = threadIdx.x, lane = tid % 64, w = tid / 64;
    const int li = lane % 16, lk = lane / 16;
    constexpr int na = 16 * NRB * KS;
    // raw factor values stay in registers until the LDS store, so their loads
    // overlap the MFMAs of the previous slice
    T ra[NA];
    double rx[NX];
    auto load = [&](int ts) {
        const int kn = min(KS, it.tn - ts);
#pragma unroll
        for (int u = 0; u < NA; u++) {
            const int q = tid + u * kMrThreads, k = q / KS, tt = q % KS;
            ra[u] = (q < na && k < it.kt && tt < kn) ? a[(size_t)k * it.ldt + it.t0 + ts + tt] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < NX; u++) {
            const int q = tid + u * kMrThreads;
            rx[u] = q / kRR < kn ? X[(size_t)(it.xoff + it.t0 + ts) * kRR + q] : 0.0;
        }
    };
    v4d acc[NRB];
#pragma unroll
    for (int ib = 0; ib < NRB; ib++) acc[ib] = v4d{0.0, 0.0, 0.0, 0.0};
    load(0);
    for (int ts = 0; ts < it.tn; ts += KS) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NA; u++) {
            const int q = tid + u * kMrThreads;
            if (q < na) as[q % KS][q / KS] = widen<T>(ra[u]);
        }
#pragma unroll
        for (int u = 0; u < NX; u++) {
            const int q = tid + u * kMrThreads;
            xs[q / kRR][q % kRR] = rx[u];
        }
        __syncthreads();
        if (ts + KS < it.tn) load(ts + KS);
#pragma unroll
        for (int kk = 0; kk < KS; kk += 4) {
            const double bf = xs[kk + lk][16 * w + li];
#pragma unroll
            for (int ib = 0; ib < NRB; ib++)
                acc[ib] = __builtin_amdgcn_mfma_f64_16x16x4f64(as[kk + lk][16 * ib + li], bf, acc[ib], 0, 0, 0);
        }
    }
#pragma unroll
    for (int ib = 0; ib < NRB; ib++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int k = 16 * ib + lk + 4 * q;
            if (k < it.kt) zpart[((size_t)it.part * 64 + k) * kRR + 16 * w + li] = acc[ib][q];
        }
}

__global__ void mr_zreduce_kernel(const MBlk1 *__restrict__ blks, int nblk, const double *__restrict__ zpart,
                                  double *__restrict__ Z, const int *__restrict__ gate)
{
    if (gate && *gate) return;
    const int b = blockIdx.x;
    if (b >= nblk) return;
    const MBlk1 B = blks[b];
    for (int q = threadIdx.x; q < B.kt * kRR; q += blockDim.x) {
        const int k = q / kRR, rr = q % kRR;
        double s = 0.0;
        for (int p = B.p0; p < B.p1; p++) s += zpart[((size_t)p * 64 + B.roff + k) * kRR + rr];
        Z[(size_t)(B.zoff + k) * kRR + rr] = s;
    }
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

// ------------------------------------------------------------ host side --
struct MrPlan {
    bool built = false;
    int kc1 = 512, cols2 = 256;    // phase-1 column chunk, phase-2 columns per item (autotuned)
    int r1[3][5][2] = {};          // phase-1 item range per (dtype, row blocks 1..4)
    int r2[3][2] = {};             // phase-2 item range per dtype
    std::vector<MItem1> it1;
    std::vector<MBlk1> blk1;
    std::vector<Item2> it2;
    std::vector<int> part0;        // ntiles + 1: partial slots of each row tile
    DevBuf<MItem1> d_it1;
    DevBuf<MBlk1> d_blk1;
    DevBuf<Item2> d_it2;
    DevBuf<int> d_part0;
    int nzpart = 0, nparts = 0;
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

// Phase-1 items.  Runs of blocks with the same column cluster and dtype are
// adjacent in the packed layout (same ldt): whole blocks are packed into row
// groups of at most 64 rows, each read against its x rows once per column
// chunk of kc1.  Ranks above 64 are split into 64-row pieces of their own.
static void mr_items1(const Plan &P, MrPlan &M, int kc1)
{
    M.it1.clear();
    M.blk1.clear();
    const auto &R = P.rows1;
    size_t i = 0;
    while (i < R.size()) {
        struct Piece { uint64_t off; int kt, zoff; };
        std::vector<Piece> grp;
        int rows = 0;
        const Row1 &f = R[i];
        if (f.kt > 64) {
            for (int k0 = 0; k0 < f.kt; k0 += 64)
                grp.push_back({f.off + (uint64_t)k0 * f.ldt * dtype_size(f.dtype), std::min(64, f.kt - k0), f.zoff + k0});
            i++;
        } else {
            size_t j = i;
            while (j < R.size() && R[j].xoff == f.xoff && R[j].nc == f.nc && R[j].dtype == f.dtype && R[j].kt <= 64 &&
                   rows + R[j].kt <= 64 &&
                   (j == i || R[j].off == R[j - 1].off + (uint64_t)R[j - 1].kt * R[j - 1].ldt * dtype_size(f.dtype))) {
                grp.push_back({R[j].off, R[j].kt, R[j].zoff});
                rows += R[j].kt;
                j++;
            }
            i = j;
        }
        // one packed run per group, except split pieces: each piece is its own run
        const bool split = f.kt > 64;
        for (size_t g0 = 0; g0 < grp.size();) {
            const size_t g1 = split ? g0 + 1 : grp.size();
            int kt = 0;
            for (size_t g = g0; g < g1; g++) kt += grp[g].kt;
            const int p0 = (int)M.it1.size();
            for (int t0 = 0; t0 < f.nc; t0 += kc1) {
                MItem1 it{};
                it.off = grp[g0].off; it.ldt = f.ldt; it.kt = kt; it.t0 = t0; it.tn = std::min(kc1, f.nc - t0);
                it.xoff = f.xoff; it.part = (int)M.it1.size(); it.dtype = f.dtype;
                M.it1.push_back(it);
            }
            const int p1 = (int)M.it1.size();
            int roff = 0;
            for (size_t g = g0; g < g1; g++) {
                M.blk1.push_back({p0, p1, grp[g].kt, grp[g].zoff, roff});
                roff += grp[g].kt;
            }
            g0 = g1;
        }
    }
    M.nzpart = (int)M.it1.size();
    // launch order: grouped by (dtype, row blocks); partial slots keep their numbers
    std::stable_sort(M.it1.begin(), M.it1.end(), [](const MItem1 &x, const MItem1 &y) {
        if (x.dtype != y.dtype) return x.dtype < y.dtype;
        return (x.kt + 15) / 16 < (y.kt + 15) / 16;
    });
    for (int d = 0; d < 3; d++)
        for (int b = 0; b <= 4; b++) M.r1[d][b][0] = M.r1[d][b][1] = 0;
    for (int q = 0; q < (int)M.it1.size(); q++) {
        const int d = M.it1[q].dtype, b = (M.it1[q].kt + 15) / 16;
        if (M.r1[d][b][1] == 0) M.r1[d][b][0] = q;
        M.r1[d][b][1] = q + 1;
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
    M.cols2 = cols2;
}

static void mr_upload(hbi_hmat *h, MrPlan &M)
{
    M.d_it1.upload(M.it1, h->stream);
    M.d_blk1.upload(M.blk1, h->stream);
    M.d_it2.upload(M.it2, h->stream);
    M.d_part0.upload(M.part0, h->stream);
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
        mr_items1(P, M, 512);
        mr_items2(h, P, M, 256);
        mr_upload(h, M);
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
static size_t zpart_cap(const MrPlan &M) { return (size_t)std::max(M.nzpart, 1) * 64 * kRR; }
static size_t ypart_cap(const hbi_hmat *h, const MrPlan &M) { return (size_t)std::max(M.nparts, 1) * h->tr * kRR; }
size_t mr_zpart_size(hbi_hmat *h, int prec) { return zpart_cap(mr_plan(h, prec)); }
size_t mr_ypart_size(hbi_hmat *h, int prec) { return ypart_cap(h, mr_plan(h, prec)); }
const int *mr_part0(hbi_hmat *h, int prec) { return mr_plan(h, prec).d_part0.p; }

// Z = a1^T X for every low-rank block (partials, then the ordered reduction)
static void phase1(const Plan &P, const MrPlan &M, const double *X, double *zpart, double *Z, hipStream_t s,
                   const int *gate)
{
    if (!M.it1.empty()) {
        for (int d = 0; d < 3; d++)
            for (int nb = 1; nb <= 4; nb++) {
                const int lo = M.r1[d][nb][0], cnt = M.r1[d][nb][1] - lo;
                if (cnt <= 0) continue;
                const MItem1 *it = M.d_it1.p + lo;
#define HBI_MR_LRVT(TT, NB) mr_lrvt_kernel<TT, NB><<<cnt, kMrThreads, 0, s>>>(it, P.buf.p, X, zpart, gate)
#define HBI_MR_LRVT_NB(TT)                         \
    switch (nb) {                                  \
    case 1: HBI_MR_LRVT(TT, 1); break;             \
    case 2: HBI_MR_LRVT(TT, 2); break;             \
    case 3: HBI_MR_LRVT(TT, 3); break;             \
    default: HBI_MR_LRVT(TT, 4); break;            \
    }
                if (d == kF64) { HBI_MR_LRVT_NB(double) }
                else if (d == kF32) { HBI_MR_LRVT_NB(float) }
                else { HBI_MR_LRVT_NB(uint16_t) }
#undef HBI_MR_LRVT_NB
#undef HBI_MR_LRVT
            }
        mr_zreduce_kernel<<<(int)M.blk1.size(), 256, 0, s>>>(M.d_blk1.p, (int)M.blk1.size(), zpart, Z, gate);
    }
}

// Y partials (per tile item) of the dense and a2 Z products
static void phase2(const hbi_hmat *h, const Plan &P, const MrPlan &M, const double *X, const double *Z, double *ypart,
                   hipStream_t s, const int *gate)
{
    for (int d = 0; d < 3; d++) {
        const int n = M.r2[d][1] - M.r2[d][0];
        if (n <= 0) continue;
        const Item2 *it = M.d_it2.p + M.r2[d][0];
        if (h->tr == 64) {
            if (d == kF64) mr_tile_kernel<double, 64><<<n, kMrThreads, 0, s>>>(it, P.d_segs.p, P.buf.p, X, Z, ypart, gate);
            else if (d == kF32) mr_tile_kernel<float, 64><<<n, kMrThreads, 0, s>>>(it, P.d_segs.p, P.buf.p, X, Z, ypart, gate);
            else mr_tile_kernel<uint16_t, 64><<<n, kMrThreads, 0, s>>>(it, P.d_segs.p, P.buf.p, X, Z, ypart, gate);
        } else {
            if (d == kF64) mr_tile_kernel<double, 32><<<n, kMrThreads, 0, s>>>(it, P.d_segs.p, P.buf.p, X, Z, ypart, gate);
            else if (d == kF32) mr_tile_kernel<float, 32><<<n, kMrThreads, 0, s>>>(it, P.d_segs.p, P.buf.p, X, Z, ypart, gate);
            else mr_tile_kernel<uint16_t, 32><<<n, kMrThreads, 0, s>>>(it, P.d_segs.p, P.buf.p, X, Z, ypart, gate);
        }
    }
}

void mr_phase1(hbi_hmat *h, int prec, const double *X, double *zpart, double *Z, hipStream_t s, const int *gate)
{
    phase1(hmat_plan(h, prec), mr_plan(h, prec), X, zpart, Z, s, gate);
}

void mr_phase2(hbi_hmat *h, int prec, const double *X, const double *Z, double *ypart, hipStream_t s, const int *gate)
{
    phase2(h, hmat_plan(h, prec), mr_plan(h, prec), X, Z, ypart, s, gate);
}

// Y partials of H X (device pointers, row-major N x 64); Y itself through mr_reduce or a consumer
void mr_apply_parts(hbi_hmat *h, int prec, const double *X, double *zpart, double *Z, double *ypart, hipStream_t s,
                    const int *gate)
{
    mr_phase1(h, prec, X, zpart, Z, s, gate);
    mr_phase2(h, prec, X, Z, ypart, s, gate);
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
    const int kc1s[] = {256, 512, 1024};
    const int c2s[] = {128, 256, 512, 1024};
    // buffers sized for the smallest chunks (most partial slots)
    mr_items1(P, M, kc1s[0]);
    mr_items2(h, P, M, c2s[0]);
    DevBuf<double> dX((size_t)h->n * kRR), dZ((size_t)std::max(h->nz, 1) * kRR);
    DevBuf<double> zp(zpart_cap(M)), yp(ypart_cap(h, M));
    HBI_HIP(hipMemsetAsync(dX.p, 0, dX.n * 8, h->stream));
    HBI_HIP(hipMemsetAsync(dZ.p, 0, dZ.n * 8, h->stream));
    auto time_it = [&](int which) {
        mr_upload(h, M);
        for (int w = 0; w < 2; w++) {
            if (which == 1) phase1(P, M, dX.p, zp.p, dZ.p, h->stream, nullptr);
            else phase2(h, P, M, dX.p, dZ.p, yp.p, h->stream, nullptr);
        }
        HBI_HIP(hipEventRecord(h->ev[0], h->stream));
        for (int r = 0; r < 5; r++) {
            if (which == 1) phase1(P, M, dX.p, zp.p, dZ.p, h->stream, nullptr);
            else phase2(h, P, M, dX.p, dZ.p, yp.p, h->stream, nullptr);
        }
        HBI_HIP(hipEventRecord(h->ev[1], h->stream));
        HBI_HIP(hipEventSynchronize(h->ev[1]));
        float ms = 0;
        HBI_HIP(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
        return ms;
    };
    int b1 = 512, b2 = 256;
    float best = 1e30f;
    for (int c : c2s) {
        mr_items2(h, P, M, c);
        const float ms = time_it(2);
        if (ms < best) { best = ms; b2 = c; }
    }
    best = 1e30f;
    for (int c : kc1s) {
        mr_items1(P, M, c);
        const float ms = time_it(1);
        if (ms < best) { best = ms; b1 = c; }
    }
    mr_items1(P, M, b1);
    mr_items2(h, P, M, b2);
    mr_upload(h, M);
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
        std::vector<double> hx((size_t)n * hbi::kRR), hy((size_t)n * hbi::kRR);
        for (int c0 = 0; c0 < nrhs; c0 += hbi::kRR) {
            const int nc = std::min(hbi::kRR, nrhs - c0);
            std::fill(hx.begin(), hx.end(), 0.0);
            for (int r = 0; r < nc; r++)
                for (int i = 0; i < n; i++) hx[(size_t)i * hbi::kRR + r] = X[(size_t)(c0 + r) * n + i];
            dX.upload(hx, h->stream);
            hbi::mr_apply_parts(h, prec, dX.p, zp.p, dZ.p, yp.p, h->stream, nullptr);
            hbi::mr_reduce(h, prec, yp.p, dY.p, h->stream);
            HBI_HIP(hipGetLastError());
            dY.download(hy.data(), hy.size(), h->stream);
            HBI_HIP(hipStreamSynchronize(h->stream));
            for (int r = 0; r < nc; r++)
                for (int i = 0; i < n; i++) Y[(size_t)(c0 + r) * n + i] = hy[(size_t)i * hbi::kRR + r];
        }
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
        HBI_HIP(hipMemsetAsync(dX.p, 0, dX.n * 8, h->stream));
        HBI_HIP(hipMemsetAsync(dZ.p, 0, dZ.n * 8, h->stream));
        for (int w = 0; w < 2; w++) hbi::mr_apply_parts(h, prec, dX.p, zp.p, dZ.p, yp.p, h->stream, nullptr);
        auto timed = [&](int which) {
            HBI_HIP(hipEventRecord(h->ev[0], h->stream));
            for (int it = 0; it < iters; it++) {
                if (which == 0) hbi::mr_phase1(h, prec, dX.p, zp.p, dZ.p, h->stream, nullptr);
                else if (which == 1) hbi::mr_phase2(h, prec, dX.p, dZ.p, yp.p, h->stream, nullptr);
                else {
                    hbi::mr_apply_parts(h, prec, dX.p, zp.p, dZ.p, yp.p, h->stream, nullptr);
                    hbi::mr_reduce(h, prec, yp.p, dY.p, h->stream);
                }
            }
            HBI_HIP(hipEventRecord(h->ev[1], h->stream));
            HBI_HIP(hipEventSynchronize(h->ev[1]));
            float t = 0;
            HBI_HIP(hipEventElapsedTime(&t, h->ev[0], h->ev[1]));
            return (double)t / iters;
        };
        (void)P;
        ms[0] = timed(0);
        ms[1] = timed(1);
        ms[2] = timed(2);
    });
}
