// hmat.hpp -- device-resident lattice H-matrix: build (GPU ACA+ / dense fill),
// packed per-precision layouts and the single- and multi-vector products.
#pragma once
#include "common.hpp"
#include "entries.hpp"
#include "partition.hpp"

namespace hbi {

enum DType : int { kF64 = 0, kF32 = 1, kBF16 = 2 };
inline int dtype_size(int d) { return d == kF64 ? 8 : (d == kF32 ? 4 : 2); }

// A run of TR x ncol column-major values inside a row tile (phase 2).
struct Seg {
    uint64_t off;   // byte offset in the plan buffer
    int ncol;       // columns
    int xoff;       // first input index (into x when src == 0, into z when src == 1)
    int src;
    int dtype;
};
// Phase-2 work item: segments [seg0, seg1) of one row tile, one dtype.
struct Item2 {
    int tile, seg0, seg1, part, dtype;
    int pad;        // 64-column FP32-core items: slices of packed segment runs (mr_slices)
};
// Low-rank work item of the single-vector product: the low-rank segments of
// one row tile, for each dtype d ncol[d] contiguous columns from byte offset
// off[d]; their z indices are zcol[zc0 ..], dtype by dtype.  In plans with
// the dense blocks in this launch (dense_b) each run starts with the tile's
// dense columns of its dtype, whose entries are x indices with bit 31 set.
struct ItemB {
    uint64_t off[3];
    int ncol[3], tile, zc0, pad;
};
// Single-vector phase 1: the a1^T rows of the low-rank blocks that share a
// column cluster and a dtype are stacked and cut into row tiles of tr rows
// (a power of two, 16 bytes of rows up to 64 rows, at most about kTile1Bytes
// per tile where the rows allow), each stored column-major (tr values per
// column, zero padded).  One work item (one workgroup) per tile.
struct Item1 {
    uint64_t off;         // byte offset of the tile
    int ncol, xoff;       // columns, x index of the first column
    int row0, nrows;      // stacked rows zrow[row0 .. row0 + nrows) -> z indices
    int tr, dtype;        // rows per column, factor dtype
};
constexpr int kTile1Bytes = 64 * 1024;
// strided FP64 -> dtype copy used to pack factors
struct CopyJob {
    long long src, dst, start;
    int which, rows, cols, src_rs, src_cs, r_lo, r_hi, c_hi, dst_rs, dst_cs, dtype, pad;
};

// a low-rank block's packed a1^T rows (phase 1 layout)
struct Row1 {
    size_t off;
    int ldt, nc, kt, xoff, zoff, dtype;
};
// consecutive segments of one tile with one dtype (phase 2 layout)
struct SegGroup {
    int tile, dtype, s0, s1;
};

struct Plan {
    bool built = false;
    std::vector<Row1> rows1;      // row-major a1^T per block (64-column product)
    std::vector<int> zrow;        // stacked row -> z index (single-vector phase 1)
    std::vector<SegGroup> groups;
    int chunk1 = 0, chunk2 = 0;   // work-item sizes in bytes (phase 2 autotuned)
    int unroll1 = 4, unroll2 = 4, unroll3 = 8;   // loads in flight per lane: phase 1, dense, low-rank (autotuned)
    bool dense_b = false;         // dense blocks streamed by launch B's row-tile items, not by launch A (autotuned)
    long long alg_dense = 0;      // dense factor bytes (in alg_bytes_p1, or p2 when dense_b)
    DevBuf<uint8_t> buf;
    size_t bytes = 0;
    std::vector<Seg> segs;
    std::vector<Item1> it1;
    std::vector<Item2> it2;        // tile items of dense blocks (launch A)
    std::vector<ItemB> itb;        // one item per tile: its low-rank blocks (launch B)
    std::vector<int> zcol;         // z index of every low-rank column of itb
    std::vector<int> tile_part0;   // ntiles + 1: the dense items' slots per tile
    DevBuf<Seg> d_segs;
    DevBuf<Item1> d_it1;
    DevBuf<int> d_zrow;
    DevBuf<Item2> d_it2;
    DevBuf<ItemB> d_itb;
    DevBuf<int> d_zcol;
    DevBuf<int> d_tile_part0;
    int nparts = 0;
    long long alg_bytes = 0;      // whole product: all factor bytes + x read + y written
    long long alg_bytes_p1 = 0;   // phase-1 kernel: a1 factors + x + z
    long long alg_bytes_p2 = 0;   // phase-2 kernel: dense + a2 factors + x + y
    int nbf16 = 0;
};

}  // namespace hbi

struct hbi_hmat {
    hipStream_t stream = nullptr;
    int n = 0, tr = 64, ntiles = 0, nz = 0, chunk_bytes = 64 * 1024;
    hbi_hmat_opts opts{};
    hbi::Partition part;
    std::vector<hbi::DevBuf<double>> geo;
    hbi::DevGeom dgeo{};
    hbi::DevBuf<int> d_lod;
    hbi::DevBuf<double> d_dense, d_lr;
    std::vector<long long> fac_off;   // per leaf: offset in d_dense (dense) or d_lr (a1 then a2)
    std::vector<int> zoff;            // per leaf: z offset (low rank), -1 otherwise
    std::vector<uint8_t> bf16_ok;     // per leaf
    hbi::Plan plan[3];
    hbi::DevBuf<double> d_x, d_y, d_z, d_ypart;
    hbi_hmat_stats st{};
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

namespace hbi {
// y = H x with device pointers on h->stream (no synchronisation)
void hmat_apply(hbi_hmat *h, int prec, const double *x, double *y);
// the product in two launches: phase 1 writes the low-rank a1^T x products
// into z (hmat_z_size doubles) and the dense blocks' tile partials into
// ypart (slots [tile_part0[t], tile_part0[t + 1]) of tile t); phase 2 adds
// the low-rank blocks' a2 z to them and writes y
size_t hmat_z_size(hbi_hmat *h, int prec);
size_t hmat_ypart_size(hbi_hmat *h, int prec);
void hmat_phase1(hbi_hmat *h, int prec, const double *x, double *z, double *ypart, hipStream_t s, const int *gate);
void hmat_phase2(hbi_hmat *h, int prec, const double *x, const double *z, const double *ypart, double *y, hipStream_t s,
                 const int *gate);
const Plan &hmat_plan(hbi_hmat *h, int prec);
// multi-RHS product (RR = 64 columns, row-major N x 64)
void mr_apply_parts(hbi_hmat *h, int prec, const double *X, float *X32, double *zpart, double *Z, double *ypart,
                    hipStream_t s, const int *gate);
void mr_reduce(hbi_hmat *h, int prec, const double *ypart, double *Y, hipStream_t s);
void mr_phase1(hbi_hmat *h, int prec, const double *X, float *X32, double *zpart, double *Z, hipStream_t s,
               const int *gate);
void mr_phase2(hbi_hmat *h, int prec, const double *X, const float *X32, const double *Z, double *ypart,
               hipStream_t s, const int *gate);
size_t mr_zpart_size(hbi_hmat *h, int prec);
size_t mr_x32_size(hbi_hmat *h, int prec);
size_t mr_ypart_size(hbi_hmat *h, int prec);
const int *mr_part0(hbi_hmat *h, int prec);
void mr_forget(const hbi_hmat *h);
}  // namespace hbi
