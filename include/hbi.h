/*
 * hbi.h -- C ABI of libhbi: the MI355X (gfx950) hot path of the HBI
 * earthquake-cycle simulator (lattice H-matrix assembly, H-matrix-vector
 * product, rate-and-state friction with adaptive Cash-Karp RK45).
 *
 * Plain C types only.  Every entry point returns an int status (HBI_OK or a
 * negative HBI_E* code) and never aborts; hbi_last_error() gives the text.
 * Vectors are in cluster order (the HACApK permutation lod), exactly as the
 * reference's st_vel%vs / st_sum%vs.  All device memory is owned by the
 * library behind opaque handles; host pointers are accepted unless a
 * function says otherwise.  One HIP stream per handle, single host thread.
 *
 * Reference interfaces replaced (amtseismo/hbi):
 *   hbi_hmat_build        HACApK_init + HACApK_generate + HACApK_construct_LH
 *                         (m_HACApK_base.f90:246, m_HACApK_use.f90:133, :263;
 *                          call sites main_LH.f90:590, 691, 693)
 *   hbi_hmat_lod          HACApK_gen_lattice_vector's lodc (m_HACApK_base.f90:3449;
 *                         call sites main_LH.f90:695-696)
 *   hbi_hmat_matvec       HACApK_adot_lattice_hyp (m_HACApK_solve.f90:220;
 *                         call sites main_LH.f90:2111, 2122, 2643)
 *   hbi_hmat_destroy      HACApK_free_leafmtxp / HACApK_finalize (main_LH.f90:1479-1485)
 *   hbi_entries           HACApK_entry_ij (m_HACApK_calc_entry_ij.f90:24)
 *   hbi_rk_derivs         derivs (main_LH.f90:2053)
 *   hbi_rk_rkck           rkck (main_LH.f90:2472)
 *   hbi_rk_step           rkqs called in the time loop (main_LH.f90:1219-1287)
 */
#ifndef HBI_H
#define HBI_H

#ifdef __cplusplus
extern "C" {
#endif

#define HBI_OK 0
#define HBI_E_ARG (-1)      /* invalid argument / configuration */
#define HBI_E_HIP (-2)      /* HIP runtime error */
#define HBI_E_NOMEM (-3)    /* device allocation failed */
#define HBI_E_STEP (-4)     /* RK step size underflow (main_LH.f90:2348-2352) */
#define HBI_E_STATE (-5)    /* object in the wrong state for this call */

/* problem kinds (st_bemv%problem) */
#define HBI_KIND_2DVS 1       /* 2-D antiplane, half-space image: '2dvs' */
#define HBI_KIND_OKADA_HS 2   /* rectangles, half-space: '3dph', '3dhr' */
#define HBI_KIND_OKADA_FS 3   /* rectangles, full space: '3dnr' */
#define HBI_KIND_TRI_FS 4     /* triangles, full space: '3dnt' */
#define HBI_KIND_TRI_HS 5     /* triangles, half-space: '3dht' (z <= 0) */
#define HBI_KIND_2DP 6        /* 2-D in-plane, full space, planar: '2dp' (matels2dp_ij) */
#define HBI_KIND_3DP 7        /* 3-D planar full space, quadrilaterals: '3dp' (matel3dp_ij) */
#define HBI_KIND_2DPH 8       /* 2-D in-plane, half-space, dipping segments: '2dph', '2dnh' (matel2dph2_ij) */
#define HBI_KIND_2DN 9        /* 2-D in-plane, full space, nonplanar segments: '2dn' (tensor2d_ij) */
#define HBI_KIND_25D 10       /* 2.5-D: rectangles of width w seen in their plane: '25d' (matelrec_ij) */
#define HBI_KIND_2DPFZ 11     /* 2-D antiplane, planar, ten image pairs at multiples of w: '2dpfz' (matels2dpFZ_ij) */
#define HBI_KIND_2DNV 12      /* 2-D in-plane, nonplanar, end-point tractions: '2dn_vector' (matel2dn_ij) */
#define HBI_KIND_2DNA 13      /* 2-D antiplane, nonplanar segments, full space: '2dna' (tensor2d3_ij) */

/* storage precision of the H-matrix factors used by a product */
#define HBI_PREC_F64 0        /* FP64 factors (the reference's arithmetic) */
#define HBI_PREC_F32 1        /* FP32 factors, FP64 accumulation */
#define HBI_PREC_MIXED 2      /* FP32 factors, BF16 where a block's ACA tolerance allows */

/* Element geometry in mesh order (st_HACApK_calc_entry,
   m_HACApK_calc_entry_ij.f90:7-19).  Unused arrays may be NULL. */
typedef struct hbi_geom {
    int kind, n;
    int vcomp;   /* 0: shear traction ('s'), 1: normal traction ('n') */
    int md;      /* slip mode (st_bemv%md): 0 's' / 'o', 1 'st', 2 'dp', 3 'open' */
    double w;    /* st_bemv%w: okada normalisation length; 25d element width */
    const double *xcol, *ycol, *zcol;
    const double *xel, *xer;
    const double *ang, *angd, *dsl, *dsd, *rake;
    const double *xs1, *xs2, *xs3, *ys1, *ys2, *ys3, *zs1, *zs2, *zs3;
    const double *ev11, *ev12, *ev13, *ev21, *ev22, *ev23, *ev31, *ev32, *ev33;
    const double *yel, *yer;   /* 2-D segment end points (2dph, 2dnh, 2dn, 25d) */
    const double *xs4, *zs4;   /* fourth corner of 3dp quadrilaterals */
    const double *ds;          /* st_bemv%ds: 2-D segment lengths (25d) */
} hbi_geom;

/* H-matrix construction options (the st_ctl%param entries that matter) */
typedef struct hbi_hmat_opts {
    int leaf;            /* param(21) leaf size; reference default 15 */
    double eta;          /* param(51) admissibility; default 2 */
    double eps_aca;      /* eps_h = param(71) */
    int kmax;            /* param(63) ACA rank cap; default 1000 */
    double bf16_margin;  /* HBI_PREC_MIXED: BF16 factors for a low-rank block when its
                            BF16 rounding error bound <= bf16_margin x its ACA tolerance;
                            0 disables BF16 */
    int tile_rows;       /* matvec row tile (32 or 64); 0 = from leaf size */
    int chunk_kb;        /* matvec work-item size target in KiB; 0 = default */
    int dense_launch;    /* single-vector product: the launch that streams the dense blocks,
                            1 = the first (with the a1^T tiles), 2 = the second (with the
                            a2 columns, per row tile); 0 = picked by the autotuner */
} hbi_hmat_opts;

typedef struct hbi_hmat_stats {
    int n, nleaf, nlr, ndense, kt_max, tile_rows;
    double kt_mean;
    long long dense_entries;     /* sum ndl*ndt over dense leaves */
    long long lr_entries;        /* sum (ndl+ndt)*kt over low-rank leaves */
    long long lr_dims;           /* sum (ndl+ndt) over low-rank leaves */
    long long stored_bytes[3];   /* packed bytes per precision (incl. padding) */
    long long alg_bytes[3];      /* algorithmic bytes per single-vector product per precision */
    long long alg_bytes_p1[3];   /* first launch: a1 and dense factors + x read + z written */
    long long alg_bytes_p2[3];   /* second launch: a2 factors + z read + y written */
    int chunk_p1[3], chunk_p2[3];   /* work-item bytes per precision chosen by the autotuner */
    int dense_launch[3];         /* launch (1 or 2) that streams the dense blocks, per precision */
    int nbf16_blocks;            /* low-rank blocks stored as BF16 in HBI_PREC_MIXED */
    int ncapped;                 /* low-rank blocks whose ACA reached kmax */
    double t_partition, t_dense, t_aca, t_pack;   /* build phase times, seconds */
} hbi_hmat_stats;

typedef struct hbi_hmat hbi_hmat;
typedef struct hbi_rk hbi_rk;

int hbi_version(void);
const char *hbi_last_error(void);
int hbi_device_count(int *count);

/* Matrix entries K(i,j) (receiver i, source j, mesh order, 0-based) evaluated
   on the GPU in FP64.  Replaces HACApK_entry_ij. */
int hbi_entries(const hbi_geom *g, int npairs, const int *i, const int *j, double *out);

/* Build the lattice H-matrix on the GPU: cluster tree and partition (host),
   dense leaves and ACA+ low-rank leaves (device, FP64), then packed
   FP64 / FP32 / mixed copies for the products. */
int hbi_hmat_build(const hbi_geom *g, const hbi_hmat_opts *opts, hbi_hmat **out);
int hbi_hmat_destroy(hbi_hmat *h);
int hbi_hmat_stats_get(const hbi_hmat *h, hbi_hmat_stats *st);
/* lod[k] = mesh element at cluster position k (0-based) */
int hbi_hmat_lod(const hbi_hmat *h, int *lod);
/* per leaf in the reference's (row, col) sorted order: ltmtx, nstrtl, ndl, nstrtt, ndt, kt */
int hbi_hmat_leaves(const hbi_hmat *h, int *out6);
/* FP64 factors of one leaf: low rank -> a1(ndt,kt) then a2(ndl,kt); dense -> a1(ndt,ndl) */
int hbi_hmat_leaf_data(const hbi_hmat *h, int ip, double *out);

/* y = H x in cluster order, with the factors of precision prec.
   on_device = 0: x, y are host pointers; 1: device pointers. */
int hbi_hmat_matvec(hbi_hmat *h, const double *x, double *y, int prec, int on_device);
/* y = H x like hbi_hmat_matvec with device pointers, but allocating nothing:
   the product's scratch (low-rank z values, dense partial slots) lives in a
   32 MiB array the library keeps on each GPU for the life of the process.
   No counterpart in the reference; for callers that cannot allocate device
   memory (code under HIP graph capture, serving loops, harnesses that forbid
   hipMalloc).  hbi_hmat_build lays out all three precisions, so the call
   only launches kernels.  A matrix whose scratch exceeds 32 MiB (its z
   values plus tr values per dense partial slot, 8 bytes each; 5 MB for the
   32768-element benchmark fault) is rejected with HBI_E_ARG.  Calls that overlap on one GPU
   share the array and race: issue them one after another.  Returns after
   the product completes, like hbi_hmat_matvec. */
int hbi_hmat_matvec_noalloc(hbi_hmat *h, const double *x, double *y, int prec);
/* Y = H X for nrhs right-hand sides, X and Y column-major N x nrhs (host).
   Computed 64 columns at a time by the multi-RHS kernels (factors read once
   per 64 columns): FP64 factors on the FP64 matrix cores; FP32 / BF16 factors
   on the FP32 matrix cores with X and the low-rank intermediates rounded to
   FP32 and the FP32 sums of at most 48 columns added in FP64.  No reference
   counterpart (the reference applies H to one vector at a time). */
int hbi_hmat_matmul(hbi_hmat *h, const double *X, double *Y, int nrhs, int prec);
/* Average device time per call over iters 64-column products on device-resident
   data, HIP events on the library stream: ms[0] = phase 1 (a1^T X of the
   low-rank blocks, with the reduction of its split rows), ms[1] = phase 2
   (row tiles), both as they run inside whole products (events between the
   phases), ms[2] = whole product (with the reduction of the row-tile
   partials), products back to back. */
int hbi_hmat_time_matmul(hbi_hmat *h, int prec, int iters, double *ms);
/* The autotuned 64-column plan of a precision (built on first use):
   info[0] phase-1 column chunk, [1] phase-2 columns per item, [2] phase-1
   items, [3] phase-2 items (= FP64 partial row tiles written), [4] phase-1
   partial slots, [5] the matrix cores used: 0 FP64, 1 FP32, 2 BF16 (split
   FP32 operands), [6] bit 0: the phase-1 items that read the same input rows
   are placed on one XCD, bits 1-3 and 4-6: register stages of the phase-1
   and phase-2 load rings (BF16 cores). */
int hbi_hmat_matmul_plan(hbi_hmat *h, int prec, int *info);
/* Measured FP64 matrix-core rate (TFLOP/s) of v_mfma_f64_16x16x4f64, the
   instruction of both 64-column kernels: waves_per_simd (1..8) waves per SIMD
   each run iters rounds of 8 independent chains, no memory traffic.  The
   ceiling the 64-column products are set against beside the data-sheet peak;
   no reference counterpart. */
int hbi_mfma_f64_rate(int waves_per_simd, int iters, double *tflops);
/* The same for v_mfma_f32_16x16x4f32, the instruction of the 64-column
   kernels for FP32 / BF16 factors (2048 flops per instruction). */
int hbi_mfma_f32_rate(int waves_per_simd, int iters, double *tflops);
/* Average device time per call of the two matvec launches over iters products
   on device-resident vectors, from HIP events on the library stream:
   ms[0] = first launch (a1^T x of the stacked low-rank row tiles, and the
   dense blocks' row tiles), ms[1] = second launch (the low-rank blocks' a2 z
   per row tile, adding the tile's dense partials into y), each bracketed by
   events; ms[2] = whole product, from iters products launched back to back
   between two events. */
int hbi_hmat_time_matvec(hbi_hmat *h, int prec, int iters, double *ms);

/* Friction / RK45 options (main_LH.f90 variables) */
typedef struct hbi_rk_opts {
    int law;          /* 0 aging, 1 slip */
    int backslip;     /* matvec input V - Vpl (1) or V (0) */
    int limitsigma;
    int prec;         /* HBI_PREC_* of the shear matvec */
    double vref, minsig, maxsig, dtmax;
    double hdiv;      /* matvec result divided by this (ds0 for okada problems), 0 = none */
    double eps_r;     /* RK relative tolerance */
    double velmin, velmax;   /* stop criteria on max V (main_LH.f90:1438-1445) */
} hbi_rk_opts;

typedef struct hbi_rk_info {
    double t, dt_next, dt_last, errmax_last;
    long long attempts, accepted, rejected;
    double vmax;       /* max |V| after the last accepted step */
    int status;        /* 0 running, 1 stopped by velmax/velmin, HBI_E_STEP on underflow,
                          HBI_RK_OUTPUT_FULL: output buffers full (drain with hbi_rk_get_output) */
} hbi_rk_info;
#define HBI_RK_OUTPUT_FULL 2

/* Per-element arrays in cluster order (host pointers). hn may be NULL (sigmaconst). */
int hbi_rk_create(hbi_hmat *hs, hbi_hmat *hn, const hbi_rk_opts *opts,
                  const double *a, const double *b, const double *dc, const double *f0,
                  const double *vpl, const double *taudot, const double *sigdot,
                  const double *area, hbi_rk **out);
int hbi_rk_destroy(hbi_rk *r);
/* y: interleaved (psi, tau, sigma) x N; slip, vel: N (may be NULL -> 0 / from y) */
int hbi_rk_set_state(hbi_rk *r, const double *y, const double *slip, const double *vel,
                     double t, double dt_next);
int hbi_rk_get_state(hbi_rk *r, double *y, double *slip, double *vel, hbi_rk_info *info);
int hbi_rk_derivs(hbi_rk *r, const double *y, double *dydx);
int hbi_rk_rkck(hbi_rk *r, const double *y, double h, double *yout, double *yerr);
/* Advance nsteps accepted steps (rkqs + post-step update) through a captured
   HIP graph of one attempt; no host synchronisation inside.  When hist is
   not NULL it receives nsteps rows of 9: t, dt, errmax, rejected attempts
   before the step, log10 max|V|, mean |slip|, mean tau/sigma (both area
   weighted), max sigma, min sigma (the monitor quantities of
   main_LH.f90:1289-1336, 1491), and vhist nsteps x nrec slip rates at
   cluster indices rec. */
int hbi_rk_step(hbi_rk *r, int nsteps, double *hist, int nrec, const int *rec, double *vhist,
                hbi_rk_info *info);
/* Advance m independent runs (e.g. distinct faults, each with its own
   H-matrix and step control) by nsteps accepted steps each: one attempt of
   every run is captured into one HIP graph (a branch per run on its own
   stream), so each lockstep round is a single graph launch and the device
   runs the faults concurrently; a run that has reached its target is gated
   off on the device.  Each run's result is bit-identical to the one
   hbi_rk_step gives.  infos (may be NULL): m entries.  No reference
   counterpart (the reference runs one fault per job). */
int hbi_rk_step_many(hbi_rk *const *rs, int m, int nsteps, hbi_rk_info *infos);
/* Device time of the last hbi_rk_step call, milliseconds (HIP events) */
int hbi_rk_last_ms(const hbi_rk *r, double *ms);

/* Opening (main_LH.f90:705-716, 2096-2143, 1279-1283): hs2, hn2 are the
   matrices of a tensile dislocation (slip mode 'open') onto shear and normal
   traction.  Every stage adds hs2 xn to the shear and hn2 xn to the normal
   stress rate (divided by hdiv like the others), xn = max(0, minsig - sigma)
   |V|; after each accepted step veln = V (max(0, minsig - sigma) + min(0,
   maxsig - sigma)) and slipn += veln dt (both start at 0 with
   hbi_rk_set_state).  Needs the normal-stress matrix (hn of hbi_rk_create);
   call before hbi_rk_set_output (snapshots then also hold veln, slipn). */
int hbi_rk_set_opening(hbi_rk *r, hbi_hmat *hs2, hbi_hmat *hn2);
int hbi_rk_get_opening(hbi_rk *r, double *veln, double *slipn);

/* Per-element local output (main_LH.f90:1099-1122, 1430-1434, output_local
   at 1495-1499; the reference's 'outloc' ids mapped to cluster order through
   lodc): hbi_rk_set_probes names up to 16 cluster-order elements (nprobe 0
   clears them).  After every accepted step of the following step calls the
   device records, for each probe p, 8 values: V, slip, tau, sigma, mu =
   tau / sigma, psi, veln, slipn (veln, slipn 0 without opening).
   hbi_rk_get_probes gives the rows of the last hbi_rk_step (or
   hbi_rk_step_many) call: *nrows accepted steps x nprobe x 8, step-major;
   rows may be NULL to query *nrows only. */
int hbi_rk_set_probes(hbi_rk *r, int nprobe, const int *idx);
int hbi_rk_get_probes(hbi_rk *r, long long *nrows, double *rows);

/* Field snapshots and the event list, decided on the device after every
   accepted step k = k0 + accepted (main_LH.f90:1339-1407):
   - event detector: an event starts when max |V| > velth (onset time, the
     mean slip and the slip field are kept, hypoloc = 1-based cluster-order
     index of the first max |V|) and ends when max |V| < velth / 2; its row is
     (count, k, onset time, Mw = (log10((mean slip - mean slip at onset)
     rigid area) + 5.9) / 1.5, hypoloc) and its coseismic slip |slip - slip
     at onset| is kept (main_LH.f90:1376-1379, output_EQslip);
   - a snapshot of vel, slip, sigma, tau (cluster order) is taken at every
     event start and end, every interval-th step and when t > tout, tout
     advancing by dtout_inter outside and dtout_co during events
     (main_LH.f90:1399-1407, output_field).
   Snapshots and finished events are held on the device; when either buffer
   fills, hbi_rk_step stops after that step with status HBI_RK_OUTPUT_FULL
   and hbi_rk_get_output must drain them before the next step.
   hbi_rk_set_state restarts the detector (no event, tout = dtout_inter, as
   main_LH.f90:1188-1195) and discards undrained output. */
typedef struct hbi_rk_out_opts {
    int interval;          /* snapshot every interval-th step; <= 0: not by count */
    int capacity;          /* snapshots held on the device (>= 1) */
    int ev_capacity;       /* finished events held on the device (>= 1) */
    int pad;
    long long k0;          /* step number before the first step of this run (restart) */
    double velth;          /* event threshold on max |V| (main_LH.f90 velth) */
    double dtout_inter;    /* snapshot spacing in time outside events (s) */
    double dtout_co;       /* snapshot spacing in time during events (s) */
} hbi_rk_out_opts;
typedef struct hbi_rk_event {
    long long count, k;    /* event number (from 1), step at which it ended */
    double onset_time;     /* t at the onset step */
    double mw;             /* moment magnitude as main_LH.f90:1376 */
    double moment;         /* mean |slip| increase over the event (area weighted) */
    int hypoloc, pad;      /* 1-based cluster-order index of max |V| at the onset */
} hbi_rk_event;
/* o == NULL switches the outputs off.  No reference counterpart as a call:
   the reference writes these files from its host loop. */
int hbi_rk_set_output(hbi_rk *r, const hbi_rk_out_opts *o);
/* Copies out and clears the held output.  nsnap, nev: counts (outputs).
   Any array may be NULL (not copied); sizes: snap_k, snap_t: capacity;
   fields: capacity x 4 x N (vel, slip, sigma, tau of each snapshot; with
   opening capacity x 6 x N, + veln, slipn); events: ev_capacity; eqslip:
   ev_capacity x N. */
int hbi_rk_get_output(hbi_rk *r, int *nsnap, long long *snap_k, double *snap_t, double *fields, int *nev,
                      hbi_rk_event *events, double *eqslip);

/* Learned step size.  No reference counterpart: the reference's only step
   controller is rkqs (main_LH.f90:2309-2357), which stays the judge of every
   step and the fallback on rejection.  A small MLP (12 -> 128 -> 128 -> 1,
   GELU; BF16 weights, FP32 biases, activations and accumulation) maps
   features of the last accepted step to a log10 correction of the step size
   h e^(-1/5) that would have put that step's error ratio e at 1.
   Parameter vector (FP32): W1 (128 x 12), b1, W2 (128 x 128), b2, W3 (128),
   b3; features are normalised as (x - mu) / sd. */
typedef struct hbi_mlp hbi_mlp;
int hbi_mlp_create(unsigned long long seed, hbi_mlp **out);
int hbi_mlp_destroy(hbi_mlp *m);
int hbi_mlp_nparams(void);
/* any pointer may be NULL (left unchanged); setting params resets Adam */
int hbi_mlp_set_params(hbi_mlp *m, const float *params, const float *mu, const float *sd);
int hbi_mlp_get_params(const hbi_mlp *m, float *params, float *mu, float *sd);
/* X: n x 12 raw features (row-major), Y: n outputs */
int hbi_mlp_forward(hbi_mlp *m, const float *X, int n, float *Y);
/* mean squared error over the n samples and its gradient (NULL: not computed) */
int hbi_mlp_loss_grad(hbi_mlp *m, const float *X, const float *T, int n, double *loss, float *grad);
/* Adam (beta 0.9 / 0.999, eps 1e-8) on shuffled minibatches; epoch_loss[e]
   = mean squared error over epoch e (may be NULL) */
int hbi_mlp_train(hbi_mlp *m, const float *X, const float *T, int n, int epochs, int batch, float lr,
                  unsigned long long seed, double *epoch_loss);

/* Step-size control of an hbi_rk: mode 0 = rkqs (the reference's), 1 =
   learned proposals after accepted steps (h = safety 10^(log10(h e^(-1/5)) +
   mlp), kept within [ratio_lo, ratio_hi] x the rkqs proposal and dtmax;
   a rejected attempt shrinks h by the rkqs rule).  The MLP's weights are
   copied; the graph is recaptured. */
int hbi_rk_set_controller(hbi_rk *r, int mode, const hbi_mlp *m, double safety, double ratio_lo, double ratio_hi);
/* Record the 12 features and (h, errmax) of up to cap accepted steps from the
   next hbi_rk_step on (cap 0: off); read them back with hbi_rk_get_trace. */
int hbi_rk_set_trace(hbi_rk *r, int cap);
int hbi_rk_get_trace(hbi_rk *r, double *feat, double *raw, int *count);

/* Ensembles: nreal <= 64 realisations of the friction parameters on one fault
   (one shear H-matrix, sigma constant or prescribed through sigdot), advanced
   in lockstep with one 64-column H product per RK stage.  Each realisation
   keeps its own step size, time, accept/reject decision (rkqs,
   main_LH.f90:2270-2370) and step count; the reference runs them as separate
   jobs.  Per-realisation arrays are realisation-major: p[r * N + i] for
   parameters, y[(r * N + i) * 3 + c] for the state; all host pointers. */
typedef struct hbi_ens hbi_ens;
int hbi_ens_create(hbi_hmat *hs, const hbi_rk_opts *opts, int nreal,
                   const double *a, const double *b, const double *dc, const double *f0,
                   const double *vpl, const double *taudot, const double *sigdot, hbi_ens **out);
int hbi_ens_destroy(hbi_ens *e);
/* Same initial time and first trial step for every realisation; slip reset to 0 */
int hbi_ens_set_state(hbi_ens *e, const double *y, double t, double dt_next);
/* Per-realisation clock: t and the next step size of each of the nreal
   realisations (call after hbi_ens_set_state, which gives every realisation
   the same t and dt_next; counters and states are kept).  With
   hbi_ens_get_clock, a run can continue from another run's states, e.g. a
   spun-up ensemble.  No counterpart in the reference (its runs restart from
   a snapshot with their own t and dt). */
int hbi_ens_set_clock(hbi_ens *e, const double *t, const double *dt_next);
int hbi_ens_get_clock(hbi_ens *e, double *t, double *dt_next);
/* Any output pointer may be NULL; t, accepted, rejected, status have nreal
   entries (status as hbi_rk_info.status) */
int hbi_ens_get_state(hbi_ens *e, double *y, double *slip, double *vel, double *t,
                      long long *accepted, long long *rejected, int *status);
/* Advance every running realisation by nsteps accepted steps (one captured
   HIP graph per lockstep attempt); *attempts = lockstep attempts launched.
   A realisation that underflows (HBI_E_STEP) or leaves [velmin, velmax]
   stops there; the others go on.  hist (may be NULL): nreal x nsteps rows of
   (t, dt, errmax, rejected attempts before the step, log10 Vmax) per accepted
   step, followed by (V, slip) of each probe element (hbi_ens_set_probes),
   zero after a stop. */
int hbi_ens_step(hbi_ens *e, int nsteps, double *hist, long long *attempts);
/* Probe elements (cluster-order indices, at most 16) whose slip rate and
   slip every history row of hbi_ens_step carries (main_LH.f90:1099-1122's
   local output, per realisation); nprobe = 0 removes them. */
int hbi_ens_set_probes(hbi_ens *e, int nprobe, const int *idx);
/* Replay a step grid (verification of reduced-precision runs against FP64 on
   a common time grid): h is nreal x cap, accepted step k (counted since
   hbi_ens_set_state) of realisation r has size h[r * cap + k]; every attempt
   is accepted (errmax is still computed and recorded), the learned proposal
   is ignored, and a step call that would run past cap steps fails.  The
   first step's size is h[r * cap], whatever dt_next was.  cap = 0 returns to
   rkqs.  No counterpart in the reference. */
int hbi_ens_set_step_grid(hbi_ens *e, long long cap, const double *h);
/* Step-size control of every realisation, as hbi_rk_set_controller (the
   features and proposals are per realisation; the MLP is shared). */
int hbi_ens_set_controller(hbi_ens *e, int mode, const hbi_mlp *m, double safety, double ratio_lo, double ratio_hi);
/* Per-realisation traces of up to cap accepted steps: feat nreal x cap x 12,
   raw nreal x cap x 2 (h, errmax), count nreal. */
int hbi_ens_set_trace(hbi_ens *e, int cap);
int hbi_ens_get_trace(hbi_ens *e, double *feat, double *raw, int *count);
int hbi_ens_last_ms(const hbi_ens *e, double *ms);

/* ---------------------------------------------------------------------------
   Fluid-driven cycle (main_fv.f90): rate-and-state friction on effective
   normal stress sigma_e = sigma - pf with permeability evolution, and
   pore-pressure diffusion along the fault solved implicitly after every
   accepted step.
   State per element (cluster order, interleaved as main_fv.f90:995-1003):
   y = (psi, tau, sigma_e, ks); pf, disp, vel alongside.  One accepted step:
     rkqs with the 4-component derivs (main_fv.f90:2039-2146, 2149-2256: the
       error is max |yerr / yout| / eps over all 4N components);
     post-RK update (1016-1028: sigma = sigma_e + pf, V from the new y);
     the implicit pore-pressure step (implicitsolver2 1727-1774, or with
       pressure_dependent implicitsolver 1678-1724, Picard iterations on the
       diffusivity) through Beuler (1776-1940): backward Euler on the 1-D
       finite-volume grid of spacing ds0 over cluster-order neighbours,
       boundary penalties (Neumann / Dirichlet), the setting's source terms,
       solved by conjugate gradients to |r|^2 < tol^2 N (tol = 1e-4, at most
       10000 iterations); the next step size is limited by dpf;
     field update (1045-1082: sigma_e = sigma - pf, disp trapezoid, V, mu, kp,
       qflow) and the stop criteria (1136-1152).
   The 1-D pore-pressure grid holds at most 8192 elements (one workgroup
   solves it, the vectors in registers and LDS).

   With imax, jmax > 0 (main_fv3dp.f90): the same cycle on a 3dp fault with
   the pore pressure on the fault's imax x jmax grid -- Beuler of
   main_fv3dp.f90:1426-1649 (finite-volume operators along j and i with their
   own boundaries, the penalties of the 'ss' / 'thrust' settings, the
   'injection' source at the centre or at wells on rate schedules), solved by
   conjugate gradients on up to 40 CUs (one launch of resident workgroups, grid-wide
   reductions) to |r|^2 < tol^2 imax jmax (tol = 1e-6, at most 1000
   iterations); implicitsolver (1314-1366) / implicitsolver2 (1368-1424) with
   dpth 0.1 on max |dpf| and the schedule's step bound; the RK step bounded
   by dtmax; the minsig rule and fields of 762-790; monitor values of
   795-799 (max V, mean disp and mu over elements: pass area = 1). */
#define HBI_FV_NEUMANN 0
#define HBI_FV_DIRICHLET 1
#define HBI_FV_INJECTIONQ 0
#define HBI_FV_INJECTIONP 1
#define HBI_FV_STEADYSTATE 2
#define HBI_FV_THRUST 3
#define HBI_FV_OTHER 4       /* 'swarm' and others: no source terms */
#define HBI_FV_MAX_CELLS 8192

typedef struct hbi_fv_opts {
    int law;                 /* 0 aging, 1 slip (evlaw) */
    int backslip;            /* matvec input V - Vpl (1) or V (0) */
    int prec;                /* HBI_PREC_* of the matvecs */
    int pressure_dependent;  /* implicitsolver (1) or implicitsolver2 (0) */
    int pfconst;             /* 1: pf is not updated */
    int bcl, bcr;            /* HBI_FV_NEUMANN / HBI_FV_DIRICHLET */
    int setting;             /* HBI_FV_INJECTIONQ ... */
    int outpertime;          /* hnext <= dtout_s (main_fv.f90:2233-2235, 1718-1720) */
    double vref, eps_r;
    double hdiv;             /* matvec results divided by this (ds0 for 3dph / 3dhr), 0 = none */
    double ds0;              /* grid spacing of the diffusion */
    double kL, kpmin, s0, eta, beta, Bs;
    double pbcl, pbcr;       /* Dirichlet boundary pressures */
    double q0, pinj;         /* source rate (steadystate / injectionq through qin) and injection pressure */
    double tinj_s;           /* injectionq: source active while t < tinj_s (seconds) */
    double dtout_s;          /* outpertime bound, seconds */
    double velmin, velmax, tmax_s;   /* stop criteria */
    /* main_fv3dp (imax, jmax > 0): the fault is a 3dp mesh of imax x jmax
       squares and the pore pressure diffuses on that 2-D grid (grid cell (i, j),
       0-based, is mesh element i jmax + j, coordinate3dp of
       main_fv3dp.f90:1174-1202); 0, 0: main_fv's 1-D grid */
    int imax, jmax;
    int bct, bcb;            /* grid: boundaries of j = 0 (top) and j = jmax - 1 (bottom); bcl / bcr are i = 0 / imax - 1 */
    double dtmax;            /* grid: hnext <= dtmax (main_fv3dp.f90:1820) */
    double minsig;           /* grid: where the RK's sigma_e < minsig, sigma = pf + minsig (main_fv3dp.f90:765-770) */
    double phi0;             /* grid: porosity of the diffusivity, penalties and sources */
} hbi_fv_opts;
#define HBI_FV_MAX_SCHEDULE 50   /* rate-schedule entries per well (main_fv3dp.f90:1107) */

typedef struct hbi_fv_info {
    double t, dt_next, dt_last, errmax_last;
    long long attempts, accepted, rejected;
    double vmax;
    int niter;               /* conjugate-gradient iterations of the last pore-pressure solve */
    int status;              /* 0 running; 1 V outside [velmin, velmax]; 3 t > tmax; 4 min sigma_e < 0.2; 5 grid mode: a
                                barrier of the pore-pressure solve timed out;
                                HBI_E_STEP on step-size underflow */
} hbi_fv_info;

typedef struct hbi_fv hbi_fv;
/* Per-element arrays (cluster order, host pointers): friction a, b, dc, f0,
   vpl; permeability kpmaxv (kpmax of the evolution law), kTv (healing time),
   phi (porosity), pfhyd (hydrostatic pressure), qin (source rates; thrust
   reads qin at 1-based index 2000 for its right boundary, so it needs N >=
   2000), area (ds, the monitor weights).  hn may be NULL (sigmaconst). */
int hbi_fv_create(hbi_hmat *hs, hbi_hmat *hn, const hbi_fv_opts *opts, const double *a, const double *b,
                  const double *dc, const double *f0, const double *vpl, const double *kpmaxv,
                  const double *kTv, const double *phi, const double *pfhyd, const double *qin,
                  const double *area, hbi_fv **out);
int hbi_fv_destroy(hbi_fv *f);
/* y: 4N as above; pf, disp: N (disp may be NULL -> 0) */
int hbi_fv_set_state(hbi_fv *f, const double *y, const double *pf, const double *disp, double t, double dt_next);
/* any output pointer may be NULL; vel = V after the last step; qflow
   (main_fv.f90:1071-1076) */
int hbi_fv_get_state(hbi_fv *f, double *y, double *pf, double *disp, double *vel, double *qflow,
                     hbi_fv_info *info);
/* dydx of main_fv.f90:2039-2146 at y (4N) */
int hbi_fv_derivs(hbi_fv *f, const double *y, double *dydx);
/* One implicit pore-pressure step of length h at time t (implicitsolver2, or
   implicitsolver with pressure_dependent) from pf, y (ks and sigma_e) and
   sigma: pfnew, the conjugate-gradient iterations and the step bound dtnxt
   (in: the proposed next step; out: after the dpf limit). */
int hbi_fv_diffuse(hbi_fv *f, const double *pf, const double *y, const double *sigma, double h, double t,
                   double *pfnew, int *niter, double *dtnxt);
/* Advance nsteps accepted steps (one captured HIP graph per attempt, the
   pore-pressure solve inside it).  hist (may be NULL): nsteps rows of 10:
   t, log10 max|V|, mean disp, mean mu, max sigma_e, min sigma_e, 1-based
   cluster index of the first max |V|, errmax, dt, CG iterations (the
   output_monitor columns of main_fv.f90:1307-1311). */
int hbi_fv_step(hbi_fv *f, int nsteps, double *hist, hbi_fv_info *info);
/* The following hbi_fv_step calls also end (status 0) after the first
   accepted step whose time exceeds t, or whose max |V| exceeds vmax_above or
   falls below vmax_below: main_fv's time-based field output (x > tout,
   main_fv.f90:1160-1163) and the event onset / end of main_fv3dp
   (max V > 1e-2, < 1e-3, main_fv3dp.f90:871-905, whose coseismic slip needs
   the fields of both steps) then fall on the last step of a call.  HUGE_VAL,
   HUGE_VAL, 0 (the defaults) disable them. */
int hbi_fv_set_break(hbi_fv *f, double t, double vmax_above, double vmax_below);
int hbi_fv_last_ms(const hbi_fv *f, double *ms);
/* Local output (main_fv.f90:862-868, output_local at 1312-1316): up to 16
   cluster-order elements; after every accepted step of the following step
   calls the device records for each 9 values: V, disp, sigma_e, tau, pf,
   sigma, mu, psi, kp.  hbi_fv_get_probes gives the last hbi_fv_step call's
   rows (*nrows x nprobe x 9, step-major; rows NULL: count only). */
int hbi_fv_set_probes(hbi_fv *f, int nprobe, const int *idx);
/* Grid mode, setting 'injection': wells on rate schedules instead of the
   centre source q0 (input_well, main_fv3dp.f90:1101-1118, used at
   1546-1565 and 1407-1419): well w at 1-based grid cell (iw[w], jw[w]) with
   kleng[w] <= HBI_FV_MAX_SCHEDULE rates qvals[w * 50 + k] at times
   qtimes[w * 50 + k] (seconds), interpolated linearly; after each step the
   first well's schedule bounds the next step at its rate changes.  nwell = 0
   returns to the centre source.  Resets the schedule position. */
int hbi_fv_set_wells(hbi_fv *f, int nwell, const int *iw, const int *jw, const int *kleng, const double *qvals,
                     const double *qtimes);
int hbi_fv_get_probes(hbi_fv *f, long long *nrows, double *rows);

#ifdef __cplusplus
}
#endif
#endif /* HBI_H */
