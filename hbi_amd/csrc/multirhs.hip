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
//            X_seg (ncol x RR); staged through LDS in 32-deep k-slices, each
//            thread owns a 4 x 4 output block (FP64 accumulation)
// Both phases reuse the packed single-vector layout of the precision's plan.
#include "hmat.hpp"

#include <algorithm>
#include <type_traits>

namespace hbi {

constexpr int kRR = 64;      // right-hand sides per product
constexpr int kKS = 32;      // k-slice staged in LDS
constexpr int kMrThreads = 256;

template <typename T> __device__ __forceinline__ double ld_as_double(const T *p, size_t i);
template <> __device__ __forceinline__ double ld_as_double<double>(const double *p, size_t i) { return p[i]; }
template <> __device__ __forceinline__ double ld_as_double<float>(const float *p, size_t i) { return (double)p[i]; }
template <> __device__ __forceinline__ double ld_as_double<uint16_t>(const uint16_t *p, size_t i)
{
    return (double)__uint_as_float(((uint32_t)p[i]) << 16);
}

// ------------------------------------------------------------- phase 1 --
struct MItem1 {
    uint64_t off;      // byte offset of the block's a1^T rows
    int ldt, kt;       // row stride (elements), rows
    int t0, tn;        // column chunk [t0, t0+tn)
    int xoff;          // first x row of the block
    int part;          // partial slot: zpart[part][kt][RR]
    int dtype;
};
struct MBlk1 {         // partial slots [p0, p1) of one block -> z rows zoff .. zoff+kt
    int p0, p1, kt, zoff;
};

// one workgroup per item: output kt x RR (kt <= 64 handled in passes of 16 rows)
template <typename T>
__global__ __launch_bounds__(kMrThreads) void mr_lrvt_kernel(const MItem1 *__restrict__ items,
                                                             const uint8_t *__restrict__ buf,
                                                             const double *__restrict__ X, double *__restrict__ zpart,
                                                             const int *__restrict__ gate)
{
    __shared__ double xs[kKS][kRR];
    __shared__ double as[16][kKS + 1];
    if (gate && *gate) return;
    const MItem1 it = items[blockIdx.x];
    const T *a = (const T *)(buf + it.off);
    const int tid = threadIdx.x;
    const int r4 = (tid % 16) * 4;      // 4 RHS
    const int k0t = tid / 16;           // one of 16 rows per pass
    for (int kb = 0; kb < it.kt; kb += 16) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int ts = 0; ts < it.tn; ts += kKS) {
            __syncthreads();
            // stage X rows t0+ts .. +kKS (kKS x RR doubles) and a1^T[kb..kb+16][ts..ts+kKS]
            for (int q = tid; q < kKS * kRR; q += kMrThreads) {
                const int tt = q / kRR, rr = q % kRR;
                const int t = ts + tt;
                xs[tt][rr] = t < it.tn ? X[(size_t)(it.xoff + it.t0 + t) * kRR + rr] : 0.0;
            }
            for (int q = tid; q < 16 * kKS; q += kMrThreads) {
                const int kk = q / kKS, tt = q % kKS;
                const int k = kb + kk, t = ts + tt;
                as[kk][tt] = (k < it.kt && t < it.tn) ? ld_as_double<T>(a, (size_t)k * it.ldt + it.t0 + t) : 0.0;
            }
            __syncthreads();
#pragma unroll 8
            for (int tt = 0; tt < kKS; tt++) {
                const double av = as[k0t][tt];
                acc[0] = fma(av, xs[tt][r4], acc[0]);
                acc[1] = fma(av, xs[tt][r4 + 1], acc[1]);
                acc[2] = fma(av, xs[tt][r4 + 2], acc[2]);
                acc[3] = fma(av, xs[tt][r4 + 3], acc[3]);
            }
        }
        const int k = kb + k0t;
        if (k < it.kt) {
            double *o = zpart + ((size_t)it.part * 64 + k) * kRR + r4;
            o[0] = acc[0]; o[1] = acc[1]; o[2] = acc[2]; o[3] = acc[3];
        }
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
        for (int p = B.p0; p < B.p1; p++) s += zpart[((size_t)p * 64 + k) * kRR + rr];
        Z[(size_t)(B.zoff + k) * kRR + rr] = s;
    }
}

// ------------------------------------------------------------- phase 2 --
template <typename T, int TR>
__global__ __launch_bounds__(kMrThreads) void mr_tile_kernel(const Item2 *__restrict__ items, const Seg *__restrict__ segs,
                                                             const uint8_t *__restrict__ buf,
                                                             const double *__restrict__ X, const double *__restrict__ Z,
                                                             double *__restrict__ ypart, const int *__restrict__ gate)
{
    // FP64 factors stay FP64 in LDS; FP32 and BF16 factors are held as FP32
    using AT = typename std::conditional<std::is_same<T, double>::value, double, float>::type;
    __shared__ double xs[kKS][kRR];
    __shared__ AT as[kKS][TR];
    if (gate && *gate) return;
    const Item2 it = items[blockIdx.x];
    const int tid = threadIdx.x;
    // thread -> (rows rb..rb+RPT-1, RHS cb..cb+3)
    constexpr int RPT = TR / 16;
    const int rb = (tid % 16) * RPT, cb = (tid / 16) * 4;
    double acc[RPT][4];
#pragma unroll
    for (int i = 0; i < RPT; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
    for (int s = it.seg0; s < it.seg1; s++) {
        const Seg sg = segs[s];
        const T *a = (const T *)(buf + sg.off);
        const double *xin = (sg.src ? Z : X) + (size_t)sg.xoff * kRR;
        for (int c0 = 0; c0 < sg.ncol; c0 += kKS) {
            __syncthreads();
            for (int q = tid; q < kKS * kRR; q += kMrThreads) {
                const int cc = q / kRR, rr = q % kRR;
                xs[cc][rr] = c0 + cc < sg.ncol ? xin[(size_t)(c0 + cc) * kRR + rr] : 0.0;
            }
            for (int q = tid; q < kKS * TR; q += kMrThreads) {
                const int cc = q / TR, r = q % TR;
                as[cc][r] = c0 + cc < sg.ncol ? (AT)ld_as_double<T>(a, (size_t)(c0 + cc) * TR + r) : (AT)0;
            }
            __syncthreads();
#pragma unroll 4
            for (int cc = 0; cc < kKS; cc++) {
                double xv[4];
#pragma unroll
                for (int j = 0; j < 4; j++) xv[j] = xs[cc][cb + j];
#pragma unroll
                for (int i = 0; i < RPT; i++) {
                    const double av = (double)as[cc][rb + i];
#pragma unroll
                    for (int j = 0; j < 4; j++) acc[i][j] = fma(av, xv[j], acc[i][j]);
                }
            }
        }
    }
    double *o = ypart + (size_t)it.part * TR * kRR;
#pragma unroll
    for (int i = 0; i < RPT; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) o[(size_t)(rb + i) * kRR + cb + j] = acc[i][j];
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
    std::vector<MItem1> it1;
    std::vector<MBlk1> blk1;
    DevBuf<MItem1> d_it1;
    DevBuf<MBlk1> d_blk1;
    int nzpart = 0;
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

const MrPlan &mr_plan(hbi_hmat *h, int prec)
{
    for (MrCache *c : g_mr)
        if (c->h == h && c->prec == prec) return c->P;
    const Plan &P = hmat_plan(h, prec);
    MrCache *c = new MrCache{h, prec, MrPlan()};
    MrPlan &M = c->P;
    const int KC = 512;   // column chunk of phase-1 items
    for (const Row1 &r : P.rows1) {
        // ranks above 64 are split into row groups of at most 64 (one z-partial slot holds 64 rows)
        for (int k0 = 0; k0 < r.kt; k0 += 64) {
            const int nk = std::min(64, r.kt - k0);
            MBlk1 b{(int)M.it1.size(), 0, nk, r.zoff + k0};
            for (int t0 = 0; t0 < r.nc; t0 += KC) {
                MItem1 it{};
                it.off = r.off + (uint64_t)k0 * r.ldt * dtype_size(r.dtype);
                it.ldt = r.ldt; it.kt = nk; it.t0 = t0; it.tn = std::min(KC, r.nc - t0);
                it.xoff = r.xoff; it.part = (int)M.it1.size(); it.dtype = r.dtype;
                M.it1.push_back(it);
            }
            b.p1 = (int)M.it1.size();
            M.blk1.push_back(b);
        }
    }
    M.nzpart = (int)M.it1.size();
    M.d_it1.upload(M.it1, h->stream);
    M.d_blk1.upload(M.blk1, h->stream);
    HBI_HIP(hipStreamSynchronize(h->stream));
    M.built = true;
    g_mr.push_back(c);
    return M;
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

size_t mr_zpart_size(hbi_hmat *h, int prec) { return (size_t)std::max(mr_plan(h, prec).nzpart, 1) * 64 * kRR; }
size_t mr_ypart_size(hbi_hmat *h, int prec) { return (size_t)std::max(hmat_plan(h, prec).nparts, 1) * h->tr * kRR; }

// Y partials of H X (device pointers, row-major N x 64); Y itself through mr_reduce or a consumer
void mr_apply_parts(hbi_hmat *h, int prec, const double *X, double *zpart, double *Z, double *ypart, hipStream_t s,
                    const int *gate)
{
    const Plan &P = hmat_plan(h, prec);
    const MrPlan &M = mr_plan(h, prec);
    if (!M.it1.empty()) {
        // items of one dtype per plan except in mixed mode: launch per dtype group
        int lo = 0;
        while (lo < (int)M.it1.size()) {
            int hi = lo;
            const int d = M.it1[lo].dtype;
            while (hi < (int)M.it1.size() && M.it1[hi].dtype == d) hi++;
            const MItem1 *it = M.d_it1.p + lo;
            if (d == kF64) mr_lrvt_kernel<double><<<hi - lo, kMrThreads, 0, s>>>(it, P.buf.p, X, zpart, gate);
            else if (d == kF32) mr_lrvt_kernel<float><<<hi - lo, kMrThreads, 0, s>>>(it, P.buf.p, X, zpart, gate);
            else mr_lrvt_kernel<uint16_t><<<hi - lo, kMrThreads, 0, s>>>(it, P.buf.p, X, zpart, gate);
            lo = hi;
        }
        mr_zreduce_kernel<<<(int)M.blk1.size(), 256, 0, s>>>(M.d_blk1.p, (int)M.blk1.size(), zpart, Z, gate);
    }
    for (int d = 0; d < 3; d++) {
        const int n = P.r2[d][1] - P.r2[d][0];
        if (n <= 0) continue;
        const Item2 *it = P.d_it2.p + P.r2[d][0];
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

void mr_reduce(hbi_hmat *h, int prec, const double *ypart, double *Y, hipStream_t s)
{
    const Plan &P = hmat_plan(h, prec);
    const size_t tot = (size_t)h->n * kRR;
    mr_tile_reduce_kernel<<<(int)((tot + 255) / 256), 256, 0, s>>>(h->n, h->tr, P.d_tile_part0.p, ypart, Y);
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
        hbi::DevBuf<double> dX((size_t)n * hbi::kRR), dY((size_t)n * hbi::kRR), dZ((size_t)(std::max(h->nz, 1)) * hbi::kRR);
        hbi::DevBuf<double> zp(hbi::mr_zpart_size(h, prec)), yp(hbi::mr_ypart_size(h, prec));
        HBI_HIP(hipMemsetAsync(dX.p, 0, dX.n * 8, h->stream));
        for (int w = 0; w < 2; w++) hbi::mr_apply_parts(h, prec, dX.p, zp.p, dZ.p, yp.p, h->stream, nullptr);
        HBI_HIP(hipEventRecord(h->ev[0], h->stream));
        for (int it = 0; it < iters; it++) {
            hbi::mr_apply_parts(h, prec, dX.p, zp.p, dZ.p, yp.p, h->stream, nullptr);
            hbi::mr_reduce(h, prec, yp.p, dY.p, h->stream);
        }
        HBI_HIP(hipEventRecord(h->ev[1], h->stream));
        HBI_HIP(hipEventSynchronize(h->ev[1]));
        float t = 0;
        HBI_HIP(hipEventElapsedTime(&t, h->ev[0], h->ev[1]));
        ms[0] = t / iters;
    });
}
