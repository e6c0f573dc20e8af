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
    int tile, seg0, seg1, part, dtype, pad;
};
// Phase-1 work item: nk rows (of ldt elements) of a1^T of one low-rank block.
struct Item1 {
    uint64_t off;
    int ldt, nc, nk, xoff, zoff, dtype;
};
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
    std::vector<Row1> rows1;
    std::vector<SegGroup> groups;
    int chunk1 = 0, chunk2 = 0;   // work-item sizes in bytes (autotuned)
    int unroll2 = 4;              // phase-2 loads in flight per lane (autotuned)
    int r1[3][2] = {}, r2[3][2] = {};   // item ranges per dtype
    DevBuf<uint8_t> buf;
    size_t bytes = 0;
    std::vector<Seg> segs;
    std::vector<Item1> it1;
    std::vector<Item2> it2;
    std::vector<int> tile_part0;   // ntiles + 1
    DevBuf<Seg> d_segs;
    DevBuf<Item1> d_it1;
    DevBuf<Item2> d_it2;
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
// the two phases separately; ypart receives tile partials (phase 2)
void hmat_phase1(hbi_hmat *h, int prec, const double *x, double *z, hipStream_t s, const int *gate);
void hmat_phase2(hbi_hmat *h, int prec, const double *x, const double *z, double *ypart, hipStream_t s,
                 const int *gate);
const Plan &hmat_plan(hbi_hmat *h, int prec);
// multi-RHS product (RR = 64 columns, row-major N x 64)
void mr_apply_parts(hbi_hmat *h, int prec, const double *X, double *zpart, double *Z, double *ypart, hipStream_t s,
                    const int *gate);
void mr_reduce(hbi_hmat *h, int prec, const double *ypart, double *Y, hipStream_t s);
void mr_phase1(hbi_hmat *h, int prec, const double *X, double *zpart, double *Z, hipStream_t s, const int *gate);
void mr_phase2(hbi_hmat *h, int prec, const double *X, const double *Z, double *ypart, hipStream_t s, const int *gate);
size_t mr_zpart_size(hbi_hmat *h, int prec);
size_t mr_ypart_size(hbi_hmat *h, int prec);
const int *mr_part0(hbi_hmat *h, int prec);
void mr_forget(const hbi_hmat *h);
}  // namespace hbi
