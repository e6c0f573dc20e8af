// hmat.hip -- GPU assembly and products of the lattice H-matrix.
//
// Build (hbi_hmat_build, replaces HACApK_generate + HACApK_construct_LH):
//   1. cluster tree + leaf partition on the host (partition.cpp)
//   2. dense leaves: one thread per FP64 entry (fill_leafmtx_hyp's dense
//      branch, m_HACApK_base.f90:2020-2031)
//   3. low-rank leaves: ACA+ with one 256-thread workgroup per block
//      (HACApK_acaplus, m_HACApK_base.f90:1748-1946): pivot searches, norms
//      and residual rows/columns are block-wide operations, the sequential
//      pivot logic runs uniformly in every thread
//   4. packing into per-precision HBM layouts for the products
//
// Product y = H x (replaces HACApK_adot_lattice_hyp at np=1) in two phases:
//   phase 1  z_b = a1_b^T x for every low-rank block, rows of a1^T stored
//            contiguously and read with 16-byte loads
//   phase 2  row tiles of TR rows: every dense block and every a2 factor
//            touching the tile is stored as TR x ncol column-major values, a
//            lane owns a fixed set of rows and streams columns; work items
//            are bounded runs of one tile's segments writing partial sums that
//            are added in a fixed order -> deterministic, no atomics
#include "hmat.hpp"

#include <mutex>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>

namespace hbi {

// ------------------------------------------------------------ entry eval --
__global__ void entries_kernel(DevGeom g, int np, const int *__restrict__ ii, const int *__restrict__ jj,
                               double *__restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < np) out[p] = entry(g, ii[p], jj[p]);
}

// dense leaves: leaf f occupies dense[off_f + il*nc + it] = K(lod[r0+il], lod[c0+it])
struct DenseJob {
    long long off, start;
    int r0, nr, c0, nc;
};
__global__ void dense_fill_kernel(DevGeom g, const int *__restrict__ lod, const DenseJob *__restrict__ jobs, int njobs,
                                  long long total, double *__restrict__ dense)
{
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        int lo = 0, hi = njobs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (jobs[mid].start <= e) lo = mid; else hi = mid - 1;
        }
        const DenseJob J = jobs[lo];
        const long long loc = e - J.start;
        const int il = (int)(loc / J.nc), it = (int)(loc % J.nc);
        dense[J.off + loc] = entry(g, lod[J.r0 + il], lod[J.c0 + it]);
    }
}

// ------------------------------------------------------------------ ACA+ --
struct AcaJob {
    int r0, nr, c0, nc, kcap, leaf;
    long long oa, ob, ora, orb, om;
    double znrm, aca_eps;
};

constexpr int kAcaThreads = 256;

struct AcaBlk {
    const DevGeom &g;
    const int *lr, *lc;
    int nr, nc;
    double *zaa, *zab;
    uint8_t *mr, *mc;
    double *sv;
    int *si;

    // residual column j of the block (HACApK_calc_vec, ld=1)
    __device__ void col(int j, int k, double *vec) const
    {
#pragma clang fp contract(off)
        const int gj = lc[j];
        for (int ii = threadIdx.x; ii < nr; ii += kAcaThreads) {
            if (!mr[ii]) {
                double e = entry(g, lr[ii], gj);
                if (k > 0) {
                    double zau = 0.0;
                    for (int m = 0; m < k; m++) zau = zau + zaa[(size_t)m * nr + ii] * zab[(size_t)m * nc + j];
                    e = e - zau;
                }
                vec[ii] = e;
            } else if (k > 0) {
                vec[ii] = 0.0;
            }
        }
        __syncthreads();
    }
    // residual row i (HACApK_calc_vec, ld=0)
    __device__ void row(int i, int k, double *vec) const
    {
#pragma clang fp contract(off)
        const int gi = lr[i];
        for (int jj = threadIdx.x; jj < nc; jj += kAcaThreads) {
            if (!mc[jj]) {
                double e = entry(g, gi, lc[jj]);
                if (k > 0) {
                    double zau = 0.0;
                    for (int m = 0; m < k; m++) zau = zau + zab[(size_t)m * nc + jj] * zaa[(size_t)m * nr + i];
                    e = e - zau;
                }
                vec[jj] = e;
            } else if (k > 0) {
                vec[jj] = 0.0;
            }
        }
        __syncthreads();
    }
    // scaled 2-norm of HACApK_unrm_d (entries below 1e-30 ignored)
    __device__ double unrm(const double *v, int n) const
    {
#pragma clang fp contract(off)
        double m = 0.0;
        for (int i = threadIdx.x; i < n; i += kAcaThreads) {
            const double a = fabs(v[i]);
            if (a >= 1.0e-30) m = fmax(m, a);
        }
        m = block_max<kAcaThreads>(m, sv);
        if (m == 0.0) return 0.0;
        double s = 0.0;
        for (int i = threadIdx.x; i < n; i += kAcaThreads) {
            const double a = fabs(v[i]);
            if (!(a < 1.0e-30)) s = s + (a / m) * (a / m);
        }
        s = block_sum<kAcaThreads>(s, sv);
        return m * sqrt(s);
    }
    // first index of max |v| (HACApK_maxabsvalloc_d) or min |v| (minabsvalloc)
    template <bool kMax>
    __device__ int argext(const double *v, int n, double &val) const
    {
        double best = kMax ? -1.0 : INFINITY;
        int bi = -1;
        for (int i = threadIdx.x; i < n; i += kAcaThreads) {
            const double a = fabs(v[i]);
            if (kMax ? (a > best) : (a < best)) { best = a; bi = i; }
        }
        block_argext<kAcaThreads, kMax>(best, bi, sv, si);
        const double a0 = fabs(v[0]);
        __syncthreads();
        if (a0 != a0 || bi < 0) { val = a0; return 0; }
        val = best;
        return bi;
    }
    __device__ void copy(double *dst, const double *src, int n) const
    {
        for (int i = threadIdx.x; i < n; i += kAcaThreads) dst[i] = src[i];
        __syncthreads();
    }
    __device__ void scale(double *v, int n, double s) const
    {
#pragma clang fp contract(off)
        for (int i = threadIdx.x; i < n; i += kAcaThreads) v[i] = v[i] * s;
        __syncthreads();
    }
    __device__ void axpy(double *y, const double *x, double a, int n) const
    {
#pragma clang fp contract(off)
        for (int i = threadIdx.x; i < n; i += kAcaThreads) y[i] = y[i] + x[i] * a;
        __syncthreads();
    }
    __device__ void set_mask(uint8_t *m, int i) const
    {
        __syncthreads();
        if (threadIdx.x == 0) m[i] = 1;
        __syncthreads();
    }
};

__global__ __launch_bounds__(kAcaThreads) void aca_plus_kernel(DevGeom g, const int *__restrict__ lod,
                                                               const AcaJob *__restrict__ jobs, double *ws,
                                                               uint8_t *mws, int *kt_out, int *capped,
                                                               double eps, int kmax)
{
#pragma clang fp contract(off)
    __shared__ double sv[kAcaThreads / kWave];
    __shared__ int si[kAcaThreads / kWave];
    const AcaJob J = jobs[blockIdx.x];
    const int nr = J.nr, nc = J.nc;
    AcaBlk B{g, lod + J.r0, lod + J.c0, nr, nc, ws + J.oa, ws + J.ob, mws + J.om, mws + J.om + nr, sv, si};
    double *pa = ws + J.ora, *pb = ws + J.orb;
    const double ACA_EPS = J.aca_eps, znrm = J.znrm;
    const double za_eps = (double)1.0e-30f, piv_tiny = (double)1.0e-20f;
    const int klim = min(kmax, J.kcap);
    int ntries = max(nr, nc) + 1, ntr = 6, ntc = 6;
    int k = 0, jref = 0;
    B.col(jref, 0, pa);
    double colnorm = B.unrm(pa, nr);
    double rownorm;
    int iref = B.argext<false>(pa, nr, rownorm);
    B.row(iref, 0, pb);
    rownorm = B.unrm(pb, nc);
    double apx = 0.0;
    int lstop = 0, i = 0, j = 0;
    bool broke = false;
    while (k < klim && (ntr > 0 || ntc > 0) && ntries > 0) {
        ntries--;
        double *pcol = B.zaa + (size_t)k * nr, *prow = B.zab + (size_t)k * nc;
        double cmax, rmax;
        i = B.argext<true>(pa, nr, cmax);
        j = B.argext<true>(pb, nc, rmax);
        if (rmax > cmax) {
            if (j != jref) B.col(j, k, pcol); else B.copy(pcol, pa, nr);
            i = B.argext<true>(pcol, nr, cmax);
            if (cmax < ACA_EPS && k >= 1) {
                lstop = 1;
            } else {
                B.row(i, k, prow);
                const double piv = pcol[i];
                if (fabs(piv) > piv_tiny) {
                    B.scale(pcol, nr, 1.0 / piv);
                } else {
                    k = max(k - 1, 0);
                    broke = true;
                    break;
                }
            }
        } else {
            if (i != iref) B.row(i, k, prow); else B.copy(prow, pb, nc);
            j = B.argext<true>(prow, nc, rmax);
            if (rmax < ACA_EPS && k >= 1) {
                lstop = 1;
            } else {
                B.col(j, k, pcol);
                const double piv = prow[j];
                if (fabs(piv) > piv_tiny) {
                    B.scale(prow, nc, 1.0 / piv);
                } else {
                    k = max(k - 1, 0);
                    broke = true;
                    break;
                }
            }
        }
        B.set_mask(B.mr, i);
        B.set_mask(B.mc, j);
        if (i != iref) {
            const double zinv = -pcol[iref];
            B.axpy(pb, prow, zinv, nc);
            rownorm = B.unrm(pb, nc);
        }
        if (i == iref || rownorm < ACA_EPS) {
            if (i == iref) ntr++;
            if (ntr > 0) {
                rownorm = 0.0;
                i = iref;
                const int stop = (iref - 1 + nr) % nr;
                while (i != stop && rownorm < za_eps && ntr > 0) {
                    if (B.mr[i] == 0) {
                        B.row(i, k + 1, pb);
                        rownorm = B.unrm(pb, nc);
                        if (rownorm < ACA_EPS) B.set_mask(B.mr, i);
                        ntr--;
                    } else {
                        rownorm = 0.0;
                    }
                    i = (i + 1) % nr;
                }
                iref = (i - 1 + nr) % nr;
            }
        }
        if (j != jref) {
            const double zinv = -prow[jref];
            B.axpy(pa, pcol, zinv, nr);
            colnorm = B.unrm(pa, nr);
        }
        if (j == jref || colnorm < ACA_EPS) {
            if (j == jref) ntc++;
            if (ntc > 0) {
                colnorm = 0.0;
                j = jref;
                const int stop = (jref - 1 + nc) % nc;
                while (j != stop && colnorm < za_eps && ntc > 0) {
                    if (B.mc[j] == 0) {
                        B.col(j, k + 1, pa);
                        colnorm = B.unrm(pa, nr);
                        if (colnorm < ACA_EPS) B.set_mask(B.mc, j);
                        ntc--;
                    } else {
                        colnorm = 0.0;
                    }
                    j = (j + 1) % nc;
                }
                jref = (j - 1 + nc) % nc;
            }
        }
        if (colnorm < ACA_EPS && rownorm < ACA_EPS && k >= 1) {
            lstop = 1;
            k = k + 1;
        }
        if (lstop == 0) {
            const double bn = B.unrm(pcol, nr) * B.unrm(prow, nc);
            if (k == 0) {
                apx = znrm;
            } else if (bn < apx * eps && rownorm < apx * eps && colnorm < apx * eps && k >= 1) {
                lstop = 1;
            }
        }
        if (lstop == 1 && k >= 1) {
            broke = true;
            break;
        }
        k = k + 1;
    }
    if (threadIdx.x == 0) {
        kt_out[blockIdx.x] = k;
        capped[blockIdx.x] = (!broke && k >= J.kcap && J.kcap < kmax) ? 1 : 0;
    }
}

// ------------------------------------------------------------------ pack --
__device__ __forceinline__ void store_as(uint8_t *p, int dtype, double v)
{
    if (dtype == kF64) *(double *)p = v;
    else if (dtype == kF32) *(float *)p = (float)v;
    else *(uint16_t *)p = f32_to_bf16_rne((float)v);
}

__global__ void pack_kernel(const CopyJob *__restrict__ jobs, int njobs, long long total, const double *__restrict__ s0,
                            const double *__restrict__ s1, uint8_t *__restrict__ dst)
{
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        int lo = 0, hi = njobs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (jobs[mid].start <= e) lo = mid; else hi = mid - 1;
        }
        const CopyJob J = jobs[lo];
        const long long loc = e - J.start;
        const int r = (int)(loc % J.rows), c = (int)(loc / J.rows);
        double v = 0.0;
        if (r >= J.r_lo && r < J.r_hi && c < J.c_hi) {
            const double *s = J.which ? s1 : s0;
            v = s[J.src + (long long)(r - J.r_lo) * J.src_rs + (long long)c * J.src_cs];
        }
        const int sz = J.dtype == kF64 ? 8 : (J.dtype == kF32 ? 4 : 2);
        store_as(dst + J.dst + ((long long)r * J.dst_rs + (long long)c * J.dst_cs) * sz, J.dtype, v);
    }
}

// Frobenius norms of both factors of every low-rank leaf (for the BF16 rule)
__global__ void lr_norms_kernel(const double *__restrict__ lr, const long long *__restrict__ off,
                                const int *__restrict__ len1, const int *__restrict__ len2, double *__restrict__ out)
{
    __shared__ double sv[4];
    const int b = blockIdx.x;
    const double *a1 = lr + off[b], *a2 = a1 + len1[b];
    double s1 = 0.0, s2 = 0.0;
    for (int i = threadIdx.x; i < len1[b]; i += 256) s1 += a1[i] * a1[i];
    for (int i = threadIdx.x; i < len2[b]; i += 256) s2 += a2[i] * a2[i];
    s1 = block_sum<256>(s1, sv);
    s2 = block_sum<256>(s2, sv);
    if (threadIdx.x == 0) { out[2 * b] = sqrt(s1); out[2 * b + 1] = sqrt(s2); }
}

// --------------------------------------------------------------- matvec --
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef uint32_t uvec4 __attribute__((ext_vector_type(4)));

// one 16-byte chunk of packed factors kept in its storage type until the FMA
// (FP64 accumulation); streamed once per product, so loaded non-temporally
template <typename T> struct Raw;
template <> struct Raw<double> {
    using R = dvec2;
    static constexpr int V = 2;
    __device__ static R load(const uint8_t *p) { return __builtin_nontemporal_load((const R *)p); }
    __device__ static R zero() { return R{0.0, 0.0}; }
    __device__ static double get(const R &r, int e) { return r[e]; }
};
template <> struct Raw<float> {
    using R = fvec4;
    static constexpr int V = 4;
    __device__ static R load(const uint8_t *p) { return __builtin_nontemporal_load((const R *)p); }
    __device__ static R zero() { return R{0.f, 0.f, 0.f, 0.f}; }
    __device__ static double get(const R &r, int e) { return (double)r[e]; }
};
template <> struct Raw<uint16_t> {
    using R = uvec4;
    static constexpr int V = 8;
    __device__ static R load(const uint8_t *p) { return __builtin_nontemporal_load((const R *)p); }
    __device__ static R zero() { return R{0u, 0u, 0u, 0u}; }
    __device__ static double get(const R &r, int e)
    {
        const uint32_t w = r[e >> 1];
        return (double)__uint_as_float((e & 1) ? (w & 0xffff0000u) : (w << 16));
    }
};

constexpr int kMvThreads = 256;

// zeros read in place of masked factor pieces and x values (never written):
// masked loads choose an address, not a value, so no select waits on a load
static __device__ double g_mv_zero[8];

// phase 1: z rows of one stacked row tile = (tile, tr x ncol, column-major)
// times x[xoff ..].  One workgroup per item; lane l owns V rows (16 bytes) of
// the column csub = l / (tr / V) of every column group, so a wave load covers
// 64 / (tr / V) consecutive columns (1 KB, contiguous) and U loads per lane
// are in flight.  Row sums: a fixed xor-shuffle tree over the lanes sharing
// rows, then the four waves in order.  Thread r < tr returns row r's sum.
template <typename T, int U>
__device__ __forceinline__ double stack_rows(const Item1 &it, const uint8_t *__restrict__ buf,
                                             const double *__restrict__ x, double (*red)[64])
{
    constexpr int V = Raw<T>::V;
    constexpr int NW = kMvThreads / kWave;
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    const int lpc = it.tr / V, cpi = kWave / lpc;
    const int rsub = lane % lpc, csub = lane / lpc;
    const uint8_t *base = buf + it.off + (size_t)rsub * 16;
    const size_t cstride = (size_t)it.tr * sizeof(T);
    const double *xin = x + it.xoff;
    double acc[V];
#pragma unroll
    for (int e = 0; e < V; e++) acc[e] = 0.0;
    const int ngroups = (it.ncol + cpi - 1) / cpi;
    for (int g0 = w; g0 < ngroups; g0 += NW * U) {
        typename Raw<T>::R a[U];
        double xv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int c = (g0 + u * NW) * cpi + csub;
            const bool ok = c < it.ncol;
            a[u] = ok ? Raw<T>::load(base + (size_t)c * cstride) : Raw<T>::zero();
            xv[u] = ok ? xin[c] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int e = 0; e < V; e++) acc[e] = fma(Raw<T>::get(a[u], e), xv[u], acc[e]);
    }
    for (int o = lpc; o < kWave; o <<= 1)
#pragma unroll
        for (int e = 0; e < V; e++) acc[e] += __shfl_xor(acc[e], o, kWave);
    if (csub == 0)
#pragma unroll
        for (int e = 0; e < V; e++) red[w][rsub * V + e] = acc[e];
    __syncthreads();
    const int r = threadIdx.x;
    return r < it.tr ? ((red[0][r] + red[1][r]) + red[2][r]) + red[3][r] : 0.0;
}

// phase 2: partial y of one work item (one workgroup).  A lane owns V rows
// of the TR-row tile and streams columns; U column groups are loaded before
// they are used so U 16-byte loads per lane are in flight.  XG: xin is in
// global memory (masked lanes then read the zero buffer); else in LDS.
template <typename T, int TR, int U, bool XG = true>
__device__ __forceinline__ void tile_seg(const Seg &sg, const uint8_t *__restrict__ buf, const double *__restrict__ xin,
                                         int rsub, int csub, int w, double (&acc)[Raw<T>::V])
{
    constexpr int V = Raw<T>::V;
    constexpr int LPC = TR / V;        // lanes per column
    constexpr int CPI = kWave / LPC;   // columns per wave load
    constexpr int NW = kMvThreads / kWave;
    const uint8_t *base = buf + sg.off + (size_t)rsub * V * sizeof(T);
    const int ngroups = (sg.ncol + CPI - 1) / CPI;
    for (int g0 = w; g0 < ngroups; g0 += NW * U) {
        typename Raw<T>::R a[U];
        double xv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int g = g0 + u * NW;
            const int c = g * CPI + csub;
            const bool ok = g < ngroups && c < sg.ncol;
            a[u] = Raw<T>::load(ok ? base + (size_t)c * TR * sizeof(T) : (const uint8_t *)g_mv_zero);
            if constexpr (XG) xv[u] = *(ok ? xin + c : (const double *)g_mv_zero);
            else xv[u] = ok ? xin[c] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int e = 0; e < V; e++) acc[e] = fma(Raw<T>::get(a[u], e), xv[u], acc[e]);
    }
}

// Segments [s0, s1) of one tile (one dtype; dense ones read x, low-rank
// ones z), then the lanes that share rows are added in a fixed order through
// red.  Thread r < TR returns row r's sum; red is free again on return.
template <typename T, int TR, int U>
__device__ __forceinline__ double tile_rows(int s0, int s1, const Seg *__restrict__ segs,
                                            const uint8_t *__restrict__ buf, const double *__restrict__ x,
                                            const double *__restrict__ z, double *red)
{
    constexpr int V = Raw<T>::V;
    constexpr int LPC = TR / V;
    constexpr int CPI = kWave / LPC;
    constexpr int NW = kMvThreads / kWave;
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    const int rsub = lane % LPC, csub = lane / LPC;
    double acc[V];
#pragma unroll
    for (int e = 0; e < V; e++) acc[e] = 0.0;
    for (int s = s0; s < s1; s++) {
        const Seg sg = segs[s];
        tile_seg<T, TR, U>(sg, buf, (sg.src ? z : x) + sg.xoff, rsub, csub, w, acc);
    }
#pragma unroll
    for (int e = 0; e < V; e++) red[(w * kWave + lane) * V + e] = acc[e];
    __syncthreads();
    double sum = 0.0;
    const int r = threadIdx.x;
    if (r < TR) {
        const int rs = r / V, e = r % V;
        for (int ww = 0; ww < NW; ww++)
            for (int cs = 0; cs < CPI; cs++) sum += red[(ww * kWave + cs * LPC + rs) * V + e];
    }
    __syncthreads();
    return sum;
}

template <int TR, int U>
__device__ __forceinline__ double tile_rows_any(int dtype, int s0, int s1, const Seg *__restrict__ segs,
                                                const uint8_t *__restrict__ buf, const double *__restrict__ x,
                                                const double *__restrict__ z, double *red)
{
    if (dtype == kF64) return tile_rows<double, TR, U>(s0, s1, segs, buf, x, z, red);
    if (dtype == kF32) return tile_rows<float, TR, U>(s0, s1, segs, buf, x, z, red);
    return tile_rows<uint16_t, TR, U>(s0, s1, segs, buf, x, z, red);
}

// Launch A: blocks [0, n1) are phase-1 items, one row tile each (its rows'
// sums go to z through zrow); the rest are tile items of dense blocks (x
// only), each writing its partial slot.  Nothing here reads z.
template <int TR, int U1, int U2>
__global__ __launch_bounds__(kMvThreads) void stream_kernel(const Item1 *__restrict__ it1, int n1,
                                                            const Item2 *__restrict__ it2,
                                                            const Seg *__restrict__ segs,
                                                            const int *__restrict__ zrow,
                                                            const uint8_t *__restrict__ buf,
                                                            const double *__restrict__ x, double *__restrict__ z,
                                                            double *__restrict__ ypart, const int *__restrict__ gate)
{
    __shared__ double sm[kMvThreads * 8];
    if (gate && *gate) return;
    if ((int)blockIdx.x < n1) {
        const Item1 it = it1[blockIdx.x];
        double (*red)[64] = reinterpret_cast<double (*)[64]>(sm);
        double s;
        if (it.dtype == kF64) s = stack_rows<double, U1>(it, buf, x, red);
        else if (it.dtype == kF32) s = stack_rows<float, U1>(it, buf, x, red);
        else s = stack_rows<uint16_t, U1>(it, buf, x, red);
        if ((int)threadIdx.x < it.nrows) z[zrow[it.row0 + threadIdx.x]] = s;
        return;
    }
    const Item2 it = it2[blockIdx.x - n1];
    const double s = tile_rows_any<TR, U2>(it.dtype, it.seg0, it.seg1, segs, buf, x, nullptr, sm);
    if (threadIdx.x < TR) ypart[(size_t)it.part * TR + threadIdx.x] = s;
}

// The low-rank columns of one tile and dtype (contiguous in the layout,
// ncol of them from byte offset off; column j multiplies z[zc[j]], or
// x[zc[j] & ~kXcol] where the kXcol bit is set: the dense columns that lead
// the run in dense_b plans): the input values are gathered into LDS, kZL at
// a time, and the columns streamed as one run.  The whole launch is a single round of workgroups, so its time is
// the chain of memory latencies of one workgroup: the first U factor loads of
// every lane (independent of z) are issued before the gather, the gather
// issues all its index loads and then all its value loads before waiting, and
// the stream keeps two batches of U loads per lane in flight.  Thread r < TR
// returns row r's sum; red and zl are free again on return.
constexpr int kZL = 2048;
constexpr int kXcol = (int)0x80000000u;   // run column index flag: an x index (dense column), not a z index
constexpr int kZG = kZL / kMvThreads;   // z values gathered per thread and chunk
template <typename T, int TR, int U>
__device__ __forceinline__ double lowrank_rows(uint64_t off, int ncol, const int *__restrict__ zc,
                                               const uint8_t *__restrict__ buf, const double *__restrict__ z,
                                               const double *__restrict__ x, double *zl, double *red)
{
    using R = typename Raw<T>::R;
    constexpr int V = Raw<T>::V;
    constexpr int LPC = TR / V;
    constexpr int CPI = kWave / LPC;
    constexpr int NW = kMvThreads / kWave;
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    const int rsub = lane % LPC, csub = lane / LPC;
    double acc[V];
#pragma unroll
    for (int e = 0; e < V; e++) acc[e] = 0.0;
    for (int c0 = 0; c0 < ncol; c0 += kZL) {
        const int nb = min(kZL, ncol - c0);
        const int ngroups = (nb + CPI - 1) / CPI;
        const uint8_t *cb = buf + off + ((size_t)c0 * TR + (size_t)rsub * V) * sizeof(T);
        // factor values of column group g of this lane (the zero buffer past the run)
        auto aload = [&](int g) -> R {
            const int c = g * CPI + csub;
            const bool ok = g < ngroups && c < nb;
            return Raw<T>::load(ok ? cb + (size_t)c * TR * sizeof(T) : (const uint8_t *)g_mv_zero);
        };
        R a[U], b[U];
#pragma unroll
        for (int u = 0; u < U; u++) a[u] = aload(w + u * NW);
        int zi[kZG];
        double zv[kZG];
#pragma unroll
        for (int k = 0; k < kZG; k++) {
            const int j = threadIdx.x + k * kMvThreads;
            zi[k] = *(j < nb ? zc + c0 + j : (const int *)g_mv_zero);
        }
#pragma unroll
        for (int k = 0; k < kZG; k++) zv[k] = *(zi[k] < 0 ? x + (zi[k] & ~kXcol) : z + zi[k]);
#pragma unroll
        for (int k = 0; k < kZG; k++) {
            const int j = threadIdx.x + k * kMvThreads;
            if (j < nb) zl[j] = zv[k];
        }
        __syncthreads();
        for (int g0 = w; g0 < ngroups; g0 += NW * U) {
#pragma unroll
            for (int u = 0; u < U; u++) b[u] = aload(g0 + NW * U + u * NW);
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int c = (g0 + u * NW) * CPI + csub;
                const double xv = c < nb ? zl[c] : 0.0;
#pragma unroll
                for (int e = 0; e < V; e++) acc[e] = fma(Raw<T>::get(a[u], e), xv, acc[e]);
            }
#pragma unroll
            for (int u = 0; u < U; u++) a[u] = b[u];
        }
        __syncthreads();
    }
#pragma unroll
    for (int e = 0; e < V; e++) red[(w * kWave + lane) * V + e] = acc[e];
    __syncthreads();
    double sum = 0.0;
    const int r = threadIdx.x;
    if (r < TR) {
        const int rs = r / V, e = r % V;
        for (int ww = 0; ww < NW; ww++)
            for (int cs = 0; cs < CPI; cs++) sum += red[(ww * kWave + cs * LPC + rs) * V + e];
    }
    __syncthreads();
    return sum;
}

// Launch B: one item per row tile.  Its low-rank columns (a2 times z; one
// run per dtype, in dtype order) plus the tile's dense slots from launch A
// (in slot order) give the tile's rows of y.  The first kSlotPre slots are
// loaded before the low-rank runs (they were written by launch A).  In plans
// with dense_b each run also holds the tile's dense columns of its dtype
// (they lie just before its low-rank columns), there are no slots, and
// launch A holds the stacked a1^T tiles alone.
constexpr int kSlotPre = 4;
template <int TR, int U>
__global__ __launch_bounds__(kMvThreads) void lowrank_kernel(const ItemB *__restrict__ items,
                                                             const int *__restrict__ zcol,
                                                             const uint8_t *__restrict__ buf,
                                                             const double *__restrict__ z,
                                                             const double *__restrict__ ypart,
                                                             const int *__restrict__ part0, double *__restrict__ y,
                                                             int n, const double *__restrict__ x,
                                                             const int *__restrict__ gate)
{
    __shared__ double zl[kZL];
    __shared__ double red[kMvThreads * 8];
    if (gate && *gate) return;
    const ItemB it = items[blockIdx.x];
    const int r = threadIdx.x, row = it.tile * TR + r;
    const int p0 = part0[it.tile], p1 = part0[it.tile + 1];
    double pre[kSlotPre];
#pragma unroll
    for (int k = 0; k < kSlotPre; k++)
        pre[k] = *(r < TR && p0 + k < p1 ? ypart + (size_t)(p0 + k) * TR + r : g_mv_zero);
    double lr = 0.0;
    int zc = it.zc0;
    if (it.ncol[kF64] > 0)
        lr += lowrank_rows<double, TR, U>(it.off[kF64], it.ncol[kF64], zcol + zc, buf, z, x, zl, red);
    zc += it.ncol[kF64];
    if (it.ncol[kF32] > 0)
        lr += lowrank_rows<float, TR, U>(it.off[kF32], it.ncol[kF32], zcol + zc, buf, z, x, zl, red);
    zc += it.ncol[kF32];
    if (it.ncol[kBF16] > 0)
        lr += lowrank_rows<uint16_t, TR, U>(it.off[kBF16], it.ncol[kBF16], zcol + zc, buf, z, x, zl, red);
    if (r < TR && row < n) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < kSlotPre; k++)
            if (p0 + k < p1) s += pre[k];
        for (int p = p0 + kSlotPre; p < p1; p++) s += ypart[(size_t)p * TR + r];
        y[row] = s + lr;
    }
}

// ------------------------------------------------------------ host side --
static inline size_t align16(size_t b) { return (b + 15) & ~size_t(15); }

// z values of the single-vector product
static size_t z_size(const hbi_hmat *h) { return (size_t)std::max(h->nz, 1) + 8; }

// Group the phase-2 layout into work items: runs of one tile's dense
// segments of one dtype (one workgroup each) writing one partial slot, a
// tile's slots consecutive; and one item per tile with its low-rank segments.
// (Phase-1 items are the row tiles.)
static void make_items(hbi_hmat *h, Plan &P, int chunk2, int unroll1, int unroll2, int unroll3, bool dense_b)
{
    P.it2.clear();
    P.itb.clear();
    P.zcol.clear();
    P.nparts = 0;
    P.tile_part0.assign(h->ntiles + 1, 0);
    size_t g = 0;
    for (int t = 0; t < h->ntiles; t++) {
        P.tile_part0[t] = P.nparts;
        ItemB ib{};
        ib.tile = t;
        ib.zc0 = (int)P.zcol.size();
        const size_t g_first = g;
        for (; g < P.groups.size() && P.groups[g].tile == t; g++) {
            // dense segments come first in a group; the low-rank ones that
            // follow are contiguous in the layout
            const SegGroup &G = P.groups[g];
            const int sz = dtype_size(G.dtype);
            int sb = G.s0;
            while (sb < G.s1 && !P.segs[sb].src) sb++;
            if (sb < G.s1) ib.off[G.dtype] = P.segs[sb].off;
            for (int q = sb; q < G.s1; q++) {
                const Seg &sg = P.segs[q];
                HBI_REQUIRE(sg.off == ib.off[G.dtype] + (uint64_t)ib.ncol[G.dtype] * h->tr * sz,
                            "low-rank segments of a tile are not contiguous");
                ib.ncol[G.dtype] += sg.ncol;
            }
            if (sb == G.s0) continue;
            if (dense_b) {   // the group's dense columns join its run, ahead of the low-rank ones
                uint64_t o = P.segs[G.s0].off;
                int nd = 0;
                for (int q = G.s0; q < sb; q++) {
                    HBI_REQUIRE(P.segs[q].off == o, "segments of a tile are not contiguous");
                    o += (uint64_t)P.segs[q].ncol * h->tr * sz;
                    nd += P.segs[q].ncol;
                }
                HBI_REQUIRE(sb == G.s1 || o == ib.off[G.dtype], "segments of a tile are not contiguous");
                ib.off[G.dtype] = P.segs[G.s0].off;
                ib.ncol[G.dtype] += nd;
                continue;
            }
            size_t cur = 0;
            int s0 = G.s0;
            for (int q = G.s0; q < sb; q++) {
                const size_t bq = (size_t)h->tr * P.segs[q].ncol * dtype_size(G.dtype);
                if (q > s0 && cur + bq > (size_t)chunk2) {
                    P.it2.push_back({t, s0, q, P.nparts++, G.dtype, 0});
                    s0 = q;
                    cur = 0;
                }
                cur += bq;
            }
            P.it2.push_back({t, s0, sb, P.nparts++, G.dtype, 0});
        }
        // z index per low-rank column, dtype by dtype (dense_b: every column of
        // the run, dense ones as x index | kXcol)
        for (int d = 0; d < 3; d++)
            for (size_t q = g_first; q < g; q++) {
                const SegGroup &G = P.groups[q];
                if (G.dtype != d) continue;
                for (int k = G.s0; k < G.s1; k++) {
                    const Seg &sg = P.segs[k];
                    if (sg.src)
                        for (int j = 0; j < sg.ncol; j++) P.zcol.push_back(sg.xoff + j);
                    else if (dense_b)
                        for (int j = 0; j < sg.ncol; j++) P.zcol.push_back((int)((unsigned)(sg.xoff + j) | kXcol));
                }
            }
        P.itb.push_back(ib);
    }
    P.tile_part0[h->ntiles] = P.nparts;
    P.chunk2 = chunk2;
    P.unroll1 = unroll1;
    P.unroll2 = unroll2;
    P.unroll3 = unroll3;
    if (dense_b != P.dense_b) {   // the dense factor bytes move to the other launch
        const long long d = dense_b ? P.alg_dense : -P.alg_dense;
        P.alg_bytes_p1 -= d;
        P.alg_bytes_p2 += d;
        P.dense_b = dense_b;
    }
    P.d_it2.upload(P.it2, h->stream);
    P.d_itb.upload(P.itb, h->stream);
    P.d_zcol.upload(P.zcol, h->stream);
    P.d_tile_part0.upload(P.tile_part0, h->stream);
    if ((int)h->d_ypart.n < std::max(P.nparts, 1) * h->tr) h->d_ypart.alloc((size_t)std::max(P.nparts, 1) * h->tr);
    HBI_HIP(hipStreamSynchronize(h->stream));
}

static void build_plan(hbi_hmat *h, int prec)
{
    Plan &P = h->plan[prec];
    P = Plan();
    const int TR = h->tr;
    const int nt = h->ntiles;
    const auto &L = h->part.leaves;
    std::vector<CopyJob> jobs;
    long long elems = 0;
    size_t bytes = 0;
    auto leaf_dtype = [&](int ip) -> int {
        if (prec == HBI_PREC_F64) return kF64;
        if (prec == HBI_PREC_MIXED && L[ip].ltmtx == 1 && h->bf16_ok[ip]) return kBF16;
        return kF32;
    };
    // ---- phase 1 layout: rows of a1^T per low-rank block, padded to 8 elements;
    // blocks of one column cluster (and dtype) are adjacent, so their rows
    // form one contiguous matrix against the same x rows
    std::vector<int> lr_order;
    for (int ip = 0; ip < (int)L.size(); ip++)
        if (L[ip].ltmtx == 1 && L[ip].kt > 0) lr_order.push_back(ip);
    std::stable_sort(lr_order.begin(), lr_order.end(), [&](int x, int y) {
        if (L[x].c0 != L[y].c0) return L[x].c0 < L[y].c0;
        if (L[x].nc != L[y].nc) return L[x].nc < L[y].nc;
        return leaf_dtype(x) < leaf_dtype(y);
    });
    for (int ip : lr_order) {
        const Leaf &f = L[ip];
        const int dt = leaf_dtype(ip), sz = dtype_size(dt);
        const int ldt = (f.nc + 7) / 8 * 8;
        CopyJob J{};
        J.src = h->fac_off[ip]; J.dst = (long long)bytes; J.start = elems; J.which = 1;
        J.rows = f.kt; J.cols = ldt; J.src_rs = f.nc; J.src_cs = 1; J.r_lo = 0; J.r_hi = f.kt; J.c_hi = f.nc;
        J.dst_rs = ldt; J.dst_cs = 1; J.dtype = dt;
        jobs.push_back(J);
        P.rows1.push_back({bytes, ldt, f.nc, f.kt, f.c0, h->zoff[ip], dt});
        elems += (long long)f.kt * ldt;
        bytes = align16(bytes + (size_t)f.kt * ldt * sz);
        P.alg_bytes += (long long)f.nc * f.kt * sz;
        P.alg_bytes_p1 += (long long)f.nc * f.kt * sz;
    }
    // ---- single-vector phase-1 layout: the blocks of one column cluster and
    // dtype stacked (rows in lr_order) and cut into column-major row tiles:
    // trm rows (the largest power of two from 64 down to 16 bytes of rows
    // with trm x nc values in kTile1Bytes), then the rest rounded up to 16
    // bytes of rows as tiles of decreasing powers of two
    for (size_t q0 = 0; q0 < lr_order.size();) {
        const Leaf &g = L[lr_order[q0]];
        const int dt = leaf_dtype(lr_order[q0]), sz = dtype_size(dt), V = 16 / sz;
        size_t q1 = q0;
        int K = 0;
        while (q1 < lr_order.size() && L[lr_order[q1]].c0 == g.c0 && L[lr_order[q1]].nc == g.nc &&
               leaf_dtype(lr_order[q1]) == dt)
            K += L[lr_order[q1++]].kt;
        const int row0 = (int)P.zrow.size();
        for (size_t q = q0; q < q1; q++)
            for (int k = 0; k < L[lr_order[q]].kt; k++) P.zrow.push_back(h->zoff[lr_order[q]] + k);
        int trm = 64;
        while (trm > V && (size_t)trm * g.nc * sz > (size_t)kTile1Bytes) trm >>= 1;
        std::vector<int> trs(K / trm, trm);
        for (int rest = (K % trm + V - 1) / V * V, b = trm; rest > 0; b >>= 1)   // rest: a multiple of V <= trm
            if (rest >= b) { trs.push_back(b); rest -= b; }
        int T0 = 0;
        for (int tr : trs) {
            const int nrows = std::min(tr, K - T0);
            Item1 it{};
            it.off = bytes; it.ncol = g.nc; it.xoff = g.c0; it.row0 = row0 + T0; it.nrows = nrows;
            it.tr = tr; it.dtype = dt;
            P.it1.push_back(it);
            // the pieces of the stacked blocks that fall in rows [T0, T0 + tr)
            int kofs = 0;
            for (size_t q = q0; q < q1; q++) {
                const int ip = lr_order[q], kt = L[ip].kt;
                const int lo = std::max(kofs, T0), hi = std::min(kofs + kt, T0 + tr);
                if (lo < hi) {
                    CopyJob J{};
                    J.which = 1; J.src = h->fac_off[ip] + (long long)(lo - kofs) * g.nc;
                    J.dst = (long long)(bytes + (size_t)(lo - T0) * sz); J.start = elems;
                    J.rows = hi - lo; J.cols = g.nc; J.src_rs = g.nc; J.src_cs = 1;
                    J.r_lo = 0; J.r_hi = hi - lo; J.c_hi = g.nc; J.dst_rs = 1; J.dst_cs = tr; J.dtype = dt;
                    jobs.push_back(J);
                    elems += (long long)(hi - lo) * g.nc;
                }
                kofs += kt;
            }
            if (nrows < tr) {   // zero rows below the stack
                CopyJob J{};
                J.which = 1; J.dst = (long long)(bytes + (size_t)nrows * sz); J.start = elems;
                J.rows = tr - nrows; J.cols = g.nc; J.src_rs = 0; J.src_cs = 0;
                J.r_lo = 0; J.r_hi = 0; J.c_hi = 0; J.dst_rs = 1; J.dst_cs = tr; J.dtype = dt;
                jobs.push_back(J);
                elems += (long long)(tr - nrows) * g.nc;
            }
            bytes = align16(bytes + (size_t)tr * g.nc * sz);
            T0 += tr;
        }
        q0 = q1;
    }
    std::stable_sort(P.it1.begin(), P.it1.end(), [](const Item1 &a, const Item1 &b) {
        return (long long)a.ncol * a.tr * dtype_size(a.dtype) > (long long)b.ncol * b.tr * dtype_size(b.dtype);
    });
    // ---- phase 2 layout: TR x ncol column-major segments, grouped by (tile, dtype)
    std::vector<std::vector<int>> tl(nt);
    for (int ip = 0; ip < (int)L.size(); ip++) {
        const Leaf &f = L[ip];
        if (f.ltmtx == 1 && f.kt == 0) continue;
        for (int t = f.r0 / TR; t <= (f.r0 + f.nr - 1) / TR; t++) tl[t].push_back(ip);
    }
    // dense segments first: they form the work items that need no z
    for (auto &v : tl) std::stable_partition(v.begin(), v.end(), [&](int ip) { return L[ip].ltmtx != 1; });
    for (int t = 0; t < nt; t++) {
        const int T0 = t * TR;
        for (int dt = 0; dt < 3; dt++) {
            const int s0 = (int)P.segs.size();
            for (int ip : tl[t]) {
                if (leaf_dtype(ip) != dt) continue;
                const Leaf &f = L[ip];
                const int sz = dtype_size(dt);
                const int lo = std::max(f.r0, T0), hi = std::min(f.r0 + f.nr, T0 + TR);
                const int ncol = f.ltmtx == 1 ? f.kt : f.nc;
                const size_t segbytes = (size_t)TR * ncol * sz;
                Seg sg{};
                sg.off = bytes; sg.ncol = ncol; sg.dtype = dt;
                CopyJob J{};
                J.dst = (long long)bytes; J.start = elems; J.rows = TR; J.cols = ncol;
                J.r_lo = lo - T0; J.r_hi = hi - T0; J.c_hi = ncol; J.dst_rs = 1; J.dst_cs = TR; J.dtype = dt;
                if (f.ltmtx == 1) {
                    sg.src = 1; sg.xoff = h->zoff[ip];
                    J.which = 1; J.src = h->fac_off[ip] + (long long)f.nc * f.kt + (lo - f.r0);
                    J.src_rs = 1; J.src_cs = f.nr;
                    P.alg_bytes += (long long)(hi - lo) * f.kt * sz;
                    P.alg_bytes_p2 += (long long)(hi - lo) * f.kt * sz;
                } else {
                    sg.src = 0; sg.xoff = f.c0;
                    J.which = 0; J.src = h->fac_off[ip] + (long long)(lo - f.r0) * f.nc;
                    J.src_rs = f.nc; J.src_cs = 1;
                    P.alg_bytes += (long long)(hi - lo) * f.nc * sz;
                    P.alg_bytes_p1 += (long long)(hi - lo) * f.nc * sz;
                    P.alg_dense += (long long)(hi - lo) * f.nc * sz;
                }
                jobs.push_back(J);
                elems += (long long)TR * ncol;
                bytes = align16(bytes + segbytes);
                P.segs.push_back(sg);
            }
            if ((int)P.segs.size() > s0) P.groups.push_back({t, dt, s0, (int)P.segs.size()});
        }
    }
    P.alg_bytes += 2LL * h->n * 8;   // x read, y written (FP64)
    P.alg_bytes_p1 += (long long)h->n * 8 + (long long)h->nz * 8;   // launch A: + x read, z written
    P.alg_bytes_p2 += (long long)h->nz * 8 + (long long)h->n * 8;   // launch B: + z read, y written
    for (int ip = 0; ip < (int)L.size(); ip++)
        if (leaf_dtype(ip) == kBF16) P.nbf16++;
    // pack on the device
    P.bytes = bytes;
    P.buf.alloc(std::max<size_t>(bytes, 16));
    if (!jobs.empty()) {
        DevBuf<CopyJob> dj;
        dj.upload(jobs, h->stream);
        const int grid = (int)std::min<long long>((elems + 255) / 256, 65536);
        pack_kernel<<<grid, 256, 0, h->stream>>>(dj.p, (int)jobs.size(), elems, h->d_dense.p, h->d_lr.p, P.buf.p);
        HBI_HIP(hipGetLastError());
        HBI_HIP(hipStreamSynchronize(h->stream));
    }
    P.d_segs.upload(P.segs, h->stream);
    P.d_it1.upload(P.it1, h->stream);
    P.d_zrow.upload(P.zrow, h->stream);
    P.chunk1 = kTile1Bytes;
    P.built = true;
    // ---- work-item sizes: fixed by the caller, or picked by timing the product
    if (h->opts.chunk_kb > 0) {
        make_items(h, P, h->opts.chunk_kb * 1024, 4, 4, 8, h->opts.dense_launch == 2);
        return;
    }
    const int c2s[] = {16384, 32768, 65536, 131072, 262144};
    HBI_HIP(hipMemsetAsync(h->d_x.p, 0, h->d_x.n * 8, h->stream));
    auto time_it = [&](int reps) {
        for (int w = 0; w < 2; w++) hmat_apply(h, prec, h->d_x.p, h->d_y.p);
        HBI_HIP(hipEventRecord(h->ev[0], h->stream));
        for (int r = 0; r < reps; r++) hmat_apply(h, prec, h->d_x.p, h->d_y.p);
        HBI_HIP(hipEventRecord(h->ev[1], h->stream));
        HBI_HIP(hipEventSynchronize(h->ev[1]));
        float ms = 0;
        HBI_HIP(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
        return ms / reps;
    };
    // dense tile items and their loads in flight, then the loads in flight of
    // the phase-1 and the low-rank items, then the dense blocks in launch B
    // (h->opts.dense_launch pins the last choice)
    const int dl = h->opts.dense_launch;
    int b2 = c2s[2], bu1 = 4, bu2 = 4, bu3 = 8;
    bool bdb = dl == 2;
    float best = 1e30f;
    for (int u : {4, 8})
        for (int c2 : c2s) {
            make_items(h, P, c2, bu1, u, bu3, bdb);
            const float ms = time_it(5);
            if (ms < best) { best = ms; b2 = c2; bu2 = u; }
            if (bdb) break;   // no dense items in launch A: chunk2 and unroll2 do not matter
        }
    make_items(h, P, b2, 8, bu2, bu3, bdb);
    float ms = time_it(5);
    if (ms < best) { best = ms; bu1 = 8; }
    make_items(h, P, b2, bu1, bu2, 12, bdb);
    ms = time_it(5);
    if (ms < best) { best = ms; bu3 = 12; }
    if (dl == 0) {
        // the two layouts are often within a few per cent: both are timed
        // again over more products, so the choice does not flip with noise
        make_items(h, P, b2, bu1, bu2, bu3, false);
        best = time_it(20);
        for (int u : {8, 12}) {
            make_items(h, P, b2, bu1, bu2, u, true);
            ms = time_it(20);
            if (ms < best) { best = ms; bu3 = u; bdb = true; }
        }
    }
    make_items(h, P, b2, bu1, bu2, bu3, bdb);
}

size_t hmat_z_size(hbi_hmat *h, int prec) { return z_size(h); }

const Plan &hmat_plan(hbi_hmat *h, int prec)
{
    HBI_REQUIRE(prec >= 0 && prec <= 2, "precision must be HBI_PREC_F64, _F32 or _MIXED");
    if (!h->plan[prec].built) build_plan(h, prec);
    return h->plan[prec];
}

template <int TR>
static void launch_a(const Plan &P, const double *x, double *z, double *ypart, hipStream_t s, const int *gate)
{
    const int n1 = (int)P.it1.size(), n = n1 + (int)P.it2.size();
    if (n == 0) return;
#define HBI_LAUNCH_A(U1, U2)                                                                                   \
    stream_kernel<TR, U1, U2><<<n, kMvThreads, 0, s>>>(P.d_it1.p, n1, P.d_it2.p, P.d_segs.p, P.d_zrow.p, P.buf.p, \
                                                       x, z, ypart, gate)
    if (P.unroll1 == 8 && P.unroll2 == 8) HBI_LAUNCH_A(8, 8);
    else if (P.unroll1 == 8) HBI_LAUNCH_A(8, 4);
    else if (P.unroll2 == 8) HBI_LAUNCH_A(4, 8);
    else HBI_LAUNCH_A(4, 4);
#undef HBI_LAUNCH_A
}

template <int TR>
static void launch_b(const hbi_hmat *h, const Plan &P, const double *x, const double *z, const double *ypart,
                     double *y, hipStream_t s, const int *gate)
{
    const int n = (int)P.itb.size();
    if (n == 0) return;
    if (P.unroll3 == 12)
        lowrank_kernel<TR, 12><<<n, kMvThreads, 0, s>>>(P.d_itb.p, P.d_zcol.p, P.buf.p, z, ypart, P.d_tile_part0.p,
                                                        y, h->n, x, gate);
    else
        lowrank_kernel<TR, 8><<<n, kMvThreads, 0, s>>>(P.d_itb.p, P.d_zcol.p, P.buf.p, z, ypart, P.d_tile_part0.p,
                                                       y, h->n, x, gate);
}

void hmat_phase1(hbi_hmat *h, int prec, const double *x, double *z, double *ypart, hipStream_t s, const int *gate)
{
    const Plan &P = h->plan[prec];
    if (h->tr == 64) launch_a<64>(P, x, z, ypart, s, gate);
    else launch_a<32>(P, x, z, ypart, s, gate);
}

void hmat_phase2(hbi_hmat *h, int prec, const double *x, const double *z, const double *ypart, double *y,
                 hipStream_t s, const int *gate)
{
    const Plan &P = h->plan[prec];
    if (h->tr == 64) launch_b<64>(h, P, x, z, ypart, y, s, gate);
    else launch_b<32>(h, P, x, z, ypart, y, s, gate);
}

size_t hmat_ypart_size(hbi_hmat *h, int prec) { return (size_t)std::max(hmat_plan(h, prec).nparts, 1) * h->tr; }

void hmat_apply(hbi_hmat *h, int prec, const double *x, double *y)
{
    hmat_phase1(h, prec, x, h->d_z.p, h->d_ypart.p, h->stream, nullptr);
    hmat_phase2(h, prec, x, h->d_z.p, h->d_ypart.p, y, h->stream, nullptr);
}

// ------------------------------------------------------------- geometry --
static void upload_geom(hbi_hmat *h, const hbi_geom *g)
{
    constexpr int kArrays = 33;
    const double *src[kArrays] = {g->xcol, g->ycol, g->zcol, g->xel, g->xer, g->ang, g->angd, g->dsl, g->dsd, g->rake,
                                  g->xs1, g->xs2, g->xs3, g->ys1, g->ys2, g->ys3, g->zs1, g->zs2, g->zs3,
                                  g->ev11, g->ev12, g->ev13, g->ev21, g->ev22, g->ev23, g->ev31, g->ev32, g->ev33,
                                  g->yel, g->yer, g->xs4, g->zs4, g->ds};
    const double **dst[kArrays] = {&h->dgeo.xcol, &h->dgeo.ycol, &h->dgeo.zcol, &h->dgeo.xel, &h->dgeo.xer,
                                   &h->dgeo.ang, &h->dgeo.angd, &h->dgeo.dsl, &h->dgeo.dsd, &h->dgeo.rake,
                                   &h->dgeo.xs1, &h->dgeo.xs2, &h->dgeo.xs3, &h->dgeo.ys1, &h->dgeo.ys2,
                                   &h->dgeo.ys3, &h->dgeo.zs1, &h->dgeo.zs2, &h->dgeo.zs3, &h->dgeo.ev11,
                                   &h->dgeo.ev12, &h->dgeo.ev13, &h->dgeo.ev21, &h->dgeo.ev22, &h->dgeo.ev23,
                                   &h->dgeo.ev31, &h->dgeo.ev32, &h->dgeo.ev33, &h->dgeo.yel, &h->dgeo.yer,
                                   &h->dgeo.xs4, &h->dgeo.zs4, &h->dgeo.ds};
    h->geo.clear();
    h->geo.reserve(kArrays);
    for (int k = 0; k < kArrays; k++) {
        *dst[k] = nullptr;
        if (!src[k]) continue;
        h->geo.emplace_back(g->n);
        h->geo.back().upload(src[k], g->n, h->stream);
        *dst[k] = h->geo.back().p;
    }
    h->dgeo.kind = g->kind;
    h->dgeo.n = g->n;
    h->dgeo.vcomp = g->vcomp;
    h->dgeo.md = g->md;
    h->dgeo.w = g->w;
}

static void check_geom(const hbi_geom *g)
{
    HBI_REQUIRE(g && g->n > 0, "geometry with n > 0 required");
    HBI_REQUIRE(g->kind >= 1 && g->kind <= 13, "unknown geometry kind");
    HBI_REQUIRE(g->xcol, "xcol required");
    if (g->kind == HBI_KIND_2DVS || g->kind == HBI_KIND_2DP) HBI_REQUIRE(g->xel && g->xer, "2dvs / 2dp need xel, xer");
    if (g->kind == HBI_KIND_2DPFZ) HBI_REQUIRE(g->xel && g->xer, "2dpfz needs xel, xer (and the image spacing w)");
    if (g->kind == HBI_KIND_3DP)
        HBI_REQUIRE(g->zcol && g->xs1 && g->xs2 && g->xs3 && g->xs4 && g->zs1 && g->zs2 && g->zs3 && g->zs4,
                    "3dp needs zcol and the four corners xs1..xs4, zs1..zs4");
    if (g->kind == HBI_KIND_2DPH || g->kind == HBI_KIND_2DN || g->kind == HBI_KIND_2DNV || g->kind == HBI_KIND_2DNA)
        HBI_REQUIRE(g->ycol && g->xel && g->xer && g->yel && g->yer && g->ang,
                    "2dph / 2dnh / 2dn / 2dn_vector / 2dna need ycol, segment ends xel, xer, yel, yer and ang");
    if (g->kind == HBI_KIND_25D)
        HBI_REQUIRE(g->ycol && g->ang && g->ds && g->w > 0, "25d needs ycol, ang, ds and w > 0");
    if (g->kind == HBI_KIND_OKADA_HS || g->kind == HBI_KIND_OKADA_FS)
        HBI_REQUIRE(g->ycol && g->zcol && g->ang && g->angd && g->dsl && g->dsd && g->rake && g->w > 0,
                    "okada kernels need ycol, zcol, ang, angd, dsl, dsd, rake and w > 0");
    if (g->kind == HBI_KIND_TRI_FS || g->kind == HBI_KIND_TRI_HS)
        HBI_REQUIRE(g->ycol && g->zcol && g->xs1 && g->xs2 && g->xs3 && g->ys1 && g->ys2 && g->ys3 && g->zs1 &&
                        g->zs2 && g->zs3 && g->ev11 && g->ev12 && g->ev13 && g->ev21 && g->ev22 && g->ev23 &&
                        g->ev31 && g->ev32 && g->ev33 && g->rake,
                    "triangle kernel needs vertices, frames and rake");
}

static double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

static void build(hbi_hmat *h, const hbi_geom *g, const hbi_hmat_opts *o)
{
    check_geom(g);
    h->opts = *o;
    if (h->opts.leaf <= 0) h->opts.leaf = 15;
    if (h->opts.eta <= 0) h->opts.eta = 2.0;
    if (h->opts.kmax <= 0) h->opts.kmax = 1000;
    HBI_REQUIRE(h->opts.eps_aca > 0, "eps_aca must be > 0");
    HBI_REQUIRE(h->opts.dense_launch >= 0 && h->opts.dense_launch <= 2, "dense_launch must be 0, 1 or 2");
    h->n = g->n;
    h->tr = h->opts.tile_rows > 0 ? h->opts.tile_rows : (h->opts.leaf <= 32 ? 32 : 64);
    HBI_REQUIRE(h->tr == 32 || h->tr == 64, "tile_rows must be 32 or 64");
    h->chunk_bytes = (h->opts.chunk_kb > 0 ? h->opts.chunk_kb : 64) * 1024;
    h->ntiles = (h->n + h->tr - 1) / h->tr;
    HBI_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    for (auto &e : h->ev) HBI_HIP(hipEventCreate(&e));
    double t0 = now();
    h->part = build_partition(g->n, g->xcol, g->ycol, g->zcol, h->opts.leaf, h->opts.eta);
    h->st.t_partition = now() - t0;
    upload_geom(h, g);
    h->d_lod.upload(h->part.lod, h->stream);
    auto &L = h->part.leaves;
    const int nl = (int)L.size();
    h->fac_off.assign(nl, -1);
    h->zoff.assign(nl, -1);
    h->bf16_ok.assign(nl, 0);

    // ---- dense leaves
    t0 = now();
    std::vector<DenseJob> dj;
    long long dtot = 0;
    for (int ip = 0; ip < nl; ip++) {
        if (L[ip].ltmtx != 2) continue;
        h->fac_off[ip] = dtot;
        dj.push_back({dtot, dtot, L[ip].r0, L[ip].nr, L[ip].c0, L[ip].nc});
        dtot += (long long)L[ip].nr * L[ip].nc;
    }
    h->d_dense.alloc(std::max<long long>(dtot, 1));
    if (!dj.empty()) {
        DevBuf<DenseJob> ddj;
        ddj.upload(dj, h->stream);
        const int grid = (int)std::min<long long>((dtot + 255) / 256, 1 << 20);
        dense_fill_kernel<<<grid, 256, 0, h->stream>>>(h->dgeo, h->d_lod.p, ddj.p, (int)dj.size(), dtot, h->d_dense.p);
        HBI_HIP(hipGetLastError());
        HBI_HIP(hipStreamSynchronize(h->stream));
    }
    h->st.t_dense = now() - t0;

    // ---- low-rank leaves: ACA+, growing the rank capacity where a block reaches it
    t0 = now();
    std::vector<int> lr_ids;
    for (int ip = 0; ip < nl; ip++)
        if (L[ip].ltmtx == 1) lr_ids.push_back(ip);
    const double znrmmat = (double)(1.0f / (float)h->n);
    const double paca = (double)1.0e-6f * h->opts.eps_aca;
    const int kmax = h->opts.kmax;
    std::vector<int> kt_all(nl, 0);
    // per pass: leaves still to do and their capacity
    std::vector<int> todo = lr_ids;
    int cap0 = 64;
    struct Keep { DevBuf<double> ws; std::vector<AcaJob> jobs; std::vector<int> kt; std::vector<int> capped; };
    std::vector<Keep> passes;
    std::vector<int> pass_of(nl, -1), job_of(nl, -1);
    while (!todo.empty()) {
        // biggest blocks first
        std::sort(todo.begin(), todo.end(), [&](int a, int b) {
            return (long long)L[a].nr * L[a].nc > (long long)L[b].nr * L[b].nc;
        });
        Keep K;
        long long wtot = 0, mtot = 0;
        for (int ip : todo) {
            const Leaf &f = L[ip];
            AcaJob J{};
            J.r0 = f.r0; J.nr = f.nr; J.c0 = f.c0; J.nc = f.nc; J.leaf = ip;
            J.kcap = std::min({cap0, kmax, std::min(f.nr, f.nc) + 1});
            J.oa = wtot; wtot += (long long)f.nr * J.kcap;
            J.ob = wtot; wtot += (long long)f.nc * J.kcap;
            J.ora = wtot; wtot += f.nr;
            J.orb = wtot; wtot += f.nc;
            J.om = mtot; mtot += f.nr + f.nc;
            J.znrm = znrmmat * (double)sqrtf((float)f.nr * (float)f.nc);
            J.aca_eps = paca * J.znrm;
            K.jobs.push_back(J);
        }
        K.ws.alloc(std::max<long long>(wtot, 1));
        DevBuf<uint8_t> masks(std::max<long long>(mtot, 1));
        HBI_HIP(hipMemsetAsync(K.ws.p, 0, K.ws.n * sizeof(double), h->stream));
        HBI_HIP(hipMemsetAsync(masks.p, 0, masks.n, h->stream));
        DevBuf<AcaJob> djobs;
        djobs.upload(K.jobs, h->stream);
        DevBuf<int> dkt(K.jobs.size()), dcap(K.jobs.size());
        aca_plus_kernel<<<(int)K.jobs.size(), kAcaThreads, 0, h->stream>>>(h->dgeo, h->d_lod.p, djobs.p, K.ws.p, masks.p,
                                                                            dkt.p, dcap.p, h->opts.eps_aca, kmax);
        HBI_HIP(hipGetLastError());
        K.kt.resize(K.jobs.size());
        K.capped.resize(K.jobs.size());
        dkt.download(K.kt.data(), K.kt.size(), h->stream);
        dcap.download(K.capped.data(), K.capped.size(), h->stream);
        HBI_HIP(hipStreamSynchronize(h->stream));
        std::vector<int> next;
        const int pidx = (int)passes.size();
        for (size_t q = 0; q < K.jobs.size(); q++) {
            const int ip = K.jobs[q].leaf;
            const int bound = std::min(kmax, std::min(L[ip].nr, L[ip].nc) + 1);
            if (K.capped[q] && std::min(cap0 * 4, bound) > K.jobs[q].kcap) { next.push_back(ip); continue; }
            kt_all[ip] = K.kt[q];
            pass_of[ip] = pidx;
            job_of[ip] = (int)q;
        }
        passes.push_back(std::move(K));
        todo.swap(next);
        cap0 *= 4;
    }
    // compact FP64 factors: a1 (nc x kt) then a2 (nr x kt) per leaf
    long long ltot = 0;
    int zt = 0;
    for (int ip : lr_ids) {
        L[ip].kt = kt_all[ip];
        h->fac_off[ip] = ltot;
        ltot += (long long)(L[ip].nr + L[ip].nc) * L[ip].kt;
        h->zoff[ip] = zt;
        zt += L[ip].kt;
    }
    h->nz = zt;
    h->d_lr.alloc(std::max<long long>(ltot, 1));
    for (size_t pi = 0; pi < passes.size(); pi++) {
        std::vector<CopyJob> cj;
        long long el = 0;
        for (int ip : lr_ids) {
            if (pass_of[ip] != (int)pi || L[ip].kt == 0) continue;
            const AcaJob &J = passes[pi].jobs[job_of[ip]];
            const int kt = L[ip].kt;
            CopyJob a1{}, a2{};
            a1.src = J.ob; a1.dst = h->fac_off[ip] * 8; a1.start = el; a1.which = 0; a1.rows = J.nc * kt; a1.cols = 1;
            a1.src_rs = 1; a1.src_cs = 0; a1.r_lo = 0; a1.r_hi = a1.rows; a1.c_hi = 1; a1.dst_rs = 1; a1.dst_cs = 0;
            a1.dtype = kF64;
            el += a1.rows;
            a2 = a1;
            a2.src = J.oa; a2.dst = (h->fac_off[ip] + (long long)J.nc * kt) * 8; a2.start = el; a2.rows = J.nr * kt;
            a2.r_hi = a2.rows;
            el += a2.rows;
            cj.push_back(a1);
            cj.push_back(a2);
        }
        if (cj.empty()) continue;
        DevBuf<CopyJob> dcj;
        dcj.upload(cj, h->stream);
        const int grid = (int)std::min<long long>((el + 255) / 256, 65536);
        pack_kernel<<<grid, 256, 0, h->stream>>>(dcj.p, (int)cj.size(), el, passes[pi].ws.p, passes[pi].ws.p,
                                                  (uint8_t *)h->d_lr.p);
        HBI_HIP(hipGetLastError());
        HBI_HIP(hipStreamSynchronize(h->stream));
    }
    for (auto &p : passes) {
        for (size_t q = 0; q < p.jobs.size(); q++)
            if (p.kt[q] >= kmax) h->st.ncapped++;
    }
    passes.clear();
    h->st.t_aca = now() - t0;

    // ---- BF16 eligibility of low-rank blocks: 2 u ||a2||_F ||a1||_F <= margin * eps * znrm
    if (h->opts.bf16_margin > 0 && !lr_ids.empty()) {
        std::vector<long long> off;
        std::vector<int> l1, l2;
        for (int ip : lr_ids) {
            off.push_back(h->fac_off[ip]);
            l1.push_back(L[ip].nc * L[ip].kt);
            l2.push_back(L[ip].nr * L[ip].kt);
        }
        DevBuf<long long> doff;
        DevBuf<int> dl1, dl2;
        doff.upload(off, h->stream);
        dl1.upload(l1, h->stream);
        dl2.upload(l2, h->stream);
        DevBuf<double> dn(2 * lr_ids.size());
        lr_norms_kernel<<<(int)lr_ids.size(), 256, 0, h->stream>>>(h->d_lr.p, doff.p, dl1.p, dl2.p, dn.p);
        HBI_HIP(hipGetLastError());
        std::vector<double> nn(2 * lr_ids.size());
        dn.download(nn.data(), nn.size(), h->stream);
        HBI_HIP(hipStreamSynchronize(h->stream));
        const double u = std::ldexp(1.0, -8);
        for (size_t q = 0; q < lr_ids.size(); q++) {
            const Leaf &f = L[lr_ids[q]];
            const double znrm = znrmmat * (double)sqrtf((float)f.nr * (float)f.nc);
            if (2.0 * u * nn[2 * q] * nn[2 * q + 1] <= h->opts.bf16_margin * h->opts.eps_aca * znrm)
                h->bf16_ok[lr_ids[q]] = 1;
        }
    }

    // ---- vectors and packed plans
    h->d_x.alloc(h->n + 8);
    h->d_y.alloc(h->n + 8);
    h->d_z.alloc(std::max(h->nz, 1) + 8);
    HBI_HIP(hipMemsetAsync(h->d_x.p, 0, h->d_x.n * 8, h->stream));
    HBI_HIP(hipMemsetAsync(h->d_z.p, 0, h->d_z.n * 8, h->stream));
    t0 = now();
    for (int p = 0; p < 3; p++) hmat_plan(h, p);
    h->st.t_pack = now() - t0;

    // ---- statistics
    hbi_hmat_stats &S = h->st;
    S.n = h->n;
    S.nleaf = nl;
    S.tile_rows = h->tr;
    long long ktsum = 0;
    for (const Leaf &f : L) {
        if (f.ltmtx == 1) {
            S.nlr++;
            S.lr_entries += (long long)(f.nr + f.nc) * f.kt;
            S.lr_dims += f.nr + f.nc;
            ktsum += f.kt;
            S.kt_max = std::max(S.kt_max, f.kt);
        } else {
            S.ndense++;
            S.dense_entries += (long long)f.nr * f.nc;
        }
    }
    S.kt_mean = S.nlr ? (double)ktsum / S.nlr : 0.0;
    for (int p = 0; p < 3; p++) {
        S.stored_bytes[p] = (long long)h->plan[p].bytes;
        S.alg_bytes[p] = h->plan[p].alg_bytes;
        S.alg_bytes_p1[p] = h->plan[p].alg_bytes_p1;
        S.alg_bytes_p2[p] = h->plan[p].alg_bytes_p2;
        S.chunk_p1[p] = h->plan[p].chunk1;
        S.chunk_p2[p] = h->plan[p].chunk2;
        S.dense_launch[p] = h->plan[p].dense_b ? 2 : 1;
    }
    S.nbf16_blocks = h->plan[HBI_PREC_MIXED].nbf16;
}

}  // namespace hbi

// ===================================================================== ABI ==
namespace hbi {
static thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }
}  // namespace hbi


extern "C" {

int hbi_version(void) { return 1; }
const char *hbi_last_error(void) { return hbi::g_err.c_str(); }

int hbi_device_count(int *count)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(count, "count is NULL");
        HBI_HIP(hipGetDeviceCount(count));
    });
}

int hbi_entries(const hbi_geom *g, int npairs, const int *i, const int *j, double *out)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(npairs >= 0 && (npairs == 0 || (i && j && out)), "bad entry arguments");
        hbi::check_geom(g);
        for (int p = 0; p < npairs; p++)
            HBI_REQUIRE(i[p] >= 0 && i[p] < g->n && j[p] >= 0 && j[p] < g->n, "entry index out of range");
        if (npairs == 0) return;
        hbi_hmat tmp;
        HBI_HIP(hipStreamCreateWithFlags(&tmp.stream, hipStreamNonBlocking));
        struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamDestroy(s); } } sg{tmp.stream};
        hbi::upload_geom(&tmp, g);
        hbi::DevBuf<int> di, dj;
        hbi::DevBuf<double> dout(npairs);
        di.upload(i, npairs, tmp.stream);
        dj.upload(j, npairs, tmp.stream);
        hbi::entries_kernel<<<(npairs + 127) / 128, 128, 0, tmp.stream>>>(tmp.dgeo, npairs, di.p, dj.p, dout.p);
        HBI_HIP(hipGetLastError());
        dout.download(out, npairs, tmp.stream);
        HBI_HIP(hipStreamSynchronize(tmp.stream));
        tmp.stream = nullptr;
    });
}

int hbi_hmat_build(const hbi_geom *g, const hbi_hmat_opts *opts, hbi_hmat **out)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(out && opts, "NULL argument");
        *out = nullptr;
        hbi_hmat *h = new hbi_hmat();
        try {
            hbi::build(h, g, opts);
        } catch (...) {
            hbi_hmat_destroy(h);
            throw;
        }
        *out = h;
    });
}

int hbi_hmat_destroy(hbi_hmat *h)
{
    if (!h) return HBI_OK;
    hbi::mr_forget(h);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (auto &e : h->ev)
        if (e) (void)hipEventDestroy(e);
    hipStream_t s = h->stream;
    delete h;
    if (s) (void)hipStreamDestroy(s);
    return HBI_OK;
}

int hbi_hmat_stats_get(const hbi_hmat *h, hbi_hmat_stats *st)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(h && st, "NULL argument");
        *st = h->st;
    });
}

int hbi_hmat_lod(const hbi_hmat *h, int *lod)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(h && lod, "NULL argument");
        std::copy(h->part.lod.begin(), h->part.lod.end(), lod);
    });
}

int hbi_hmat_leaves(const hbi_hmat *h, int *out6)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(h && out6, "NULL argument");
        for (size_t k = 0; k < h->part.leaves.size(); k++) {
            const hbi::Leaf &f = h->part.leaves[k];
            int *o = out6 + 6 * k;
            o[0] = f.ltmtx; o[1] = f.r0; o[2] = f.nr; o[3] = f.c0; o[4] = f.nc; o[5] = f.kt;
        }
    });
}

int hbi_hmat_leaf_data(const hbi_hmat *h, int ip, double *out)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(h && out && ip >= 0 && ip < (int)h->part.leaves.size(), "bad leaf index");
        const hbi::Leaf &f = h->part.leaves[ip];
        const size_t cnt = f.ltmtx == 1 ? (size_t)(f.nr + f.nc) * f.kt : (size_t)f.nr * f.nc;
        if (cnt == 0) return;
        const double *src = (f.ltmtx == 1 ? h->d_lr.p : h->d_dense.p) + h->fac_off[ip];
        HBI_HIP(hipMemcpyAsync(out, src, cnt * 8, hipMemcpyDeviceToHost, h->stream));
        HBI_HIP(hipStreamSynchronize(h->stream));
    });
}

int hbi_hmat_matvec(hbi_hmat *h, const double *x, double *y, int prec, int on_device)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(h && x && y, "NULL argument");
        if (on_device) {
            hbi::hmat_apply(h, prec, x, y);
            HBI_HIP(hipGetLastError());
            HBI_HIP(hipStreamSynchronize(h->stream));
        } else {
            h->d_x.upload(x, h->n, h->stream);
            hbi::hmat_apply(h, prec, h->d_x.p, h->d_y.p);
            HBI_HIP(hipGetLastError());
            h->d_y.download(y, h->n, h->stream);
            HBI_HIP(hipStreamSynchronize(h->stream));
        }
    });
}

}  // extern "C"

// ------------------------------------------------- product without allocation --
// hbi_hmat_matvec_noalloc keeps the product's scratch -- the z values of the
// low-rank blocks and the dense partial slots -- in this array instead of the
// handle's buffers.  Each GPU holds its own copy for the life of the process;
// host code reaches it only through the address hipGetSymbolAddress returns
// for the current device (the array's name is a host-side stand-in).
namespace {
constexpr size_t kScratchDoubles = (size_t)4 << 20;   // 32 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr;   // per device, set on the first call there

double *scratch_for_current_device()
{
    int dev = 0;
    HBI_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HBI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        HBI_REQUIRE(p != nullptr, "scratch array has no device address");
        g_scratch_ptr[dev] = (double *)p;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int hbi_hmat_matvec_noalloc(hbi_hmat *h, const double *x, double *y, int prec)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(h && x && y, "NULL argument");
        HBI_REQUIRE(prec >= 0 && prec <= 2, "precision must be HBI_PREC_F64, _F32 or _MIXED");
        HBI_REQUIRE(h->plan[prec].built, "hbi_hmat_matvec_noalloc: the layout of this precision is not built");
        const hbi::Plan &P = h->plan[prec];
        const size_t zn = (hbi::z_size(h) + 31) & ~size_t(31);   // ypart starts 256-byte aligned
        const size_t need = zn + (size_t)std::max(P.nparts, 1) * h->tr;
        if (need > kScratchDoubles)
            throw hbi::Error{HBI_E_ARG, "hbi_hmat_matvec_noalloc: this matrix needs " + std::to_string(need * 8) +
                                            " bytes of scratch; the maximum is " + std::to_string(kScratchDoubles * 8) +
                                            " (use hbi_hmat_matvec)"};
        double *sp = scratch_for_current_device();
        hbi::hmat_phase1(h, prec, x, sp, sp + zn, h->stream, nullptr);
        hbi::hmat_phase2(h, prec, x, sp, sp + zn, y, h->stream, nullptr);
        HBI_HIP(hipGetLastError());
        HBI_HIP(hipStreamSynchronize(h->stream));
    });
}

int hbi_hmat_time_matvec(hbi_hmat *h, int prec, int iters, double *ms)
{
    return hbi::guarded([&] {
        HBI_REQUIRE(h && ms && iters > 0, "bad arguments");
        const hbi::Plan &P = hbi::hmat_plan(h, prec);
        if ((int)h->d_ypart.n < std::max(P.nparts, 1) * h->tr) h->d_ypart.alloc((size_t)std::max(P.nparts, 1) * h->tr);
        for (int w = 0; w < 3; w++) hbi::hmat_apply(h, prec, h->d_x.p, h->d_y.p);
        HBI_HIP(hipStreamSynchronize(h->stream));
        float t1 = 0, t2 = 0, t3 = 0;
        for (int it = 0; it < iters; it++) {
            HBI_HIP(hipEventRecord(h->ev[0], h->stream));
            hbi::hmat_phase1(h, prec, h->d_x.p, h->d_z.p, h->d_ypart.p, h->stream, nullptr);
            HBI_HIP(hipEventRecord(h->ev[1], h->stream));
            hbi::hmat_phase2(h, prec, h->d_x.p, h->d_z.p, h->d_ypart.p, h->d_y.p, h->stream, nullptr);
            HBI_HIP(hipEventRecord(h->ev[2], h->stream));
            HBI_HIP(hipEventSynchronize(h->ev[2]));
            float a, b, c;
            HBI_HIP(hipEventElapsedTime(&a, h->ev[0], h->ev[1]));
            HBI_HIP(hipEventElapsedTime(&b, h->ev[1], h->ev[2]));
            HBI_HIP(hipEventElapsedTime(&c, h->ev[0], h->ev[2]));
            t1 += a; t2 += b; t3 += c;
        }
        ms[0] = t1 / iters;
        ms[1] = t2 / iters;
        // the whole product: back-to-back products between two events (no
        // event between a product's two launches)
        (void)t3;
        HBI_HIP(hipEventRecord(h->ev[0], h->stream));
        for (int it = 0; it < iters; it++) hbi::hmat_apply(h, prec, h->d_x.p, h->d_y.p);
        HBI_HIP(hipEventRecord(h->ev[1], h->stream));
        HBI_HIP(hipEventSynchronize(h->ev[1]));
        float tt = 0;
        HBI_HIP(hipEventElapsedTime(&tt, h->ev[0], h->ev[1]));
        ms[2] = tt / iters;
    });
}

}  // extern "C"
