/*
 * graspci.h -- C ABI of the MI355X rci90 engine (libgraspci.so).
 *
 * Plain C types only: pointers, sizes, int status codes (0 = ok).  Every
 * entry point can be bound from Fortran through ISO_C_BINDING (see
 * grasp_amd/fortran/gci_bind.f90) and from Python through ctypes.  Unless a
 * name ends in _device, array arguments are host memory.
 *
 * Index conventions follow the Fortran callers: orbital labels and matrix
 * column indices are 1-based unless an index_base argument says otherwise;
 * row pointers are int64 cumulative row ends (IENDC with IENDC(0) = 0).
 */
#ifndef GRASPCI_H
#define GRASPCI_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define GCI_OK 0
#define GCI_ERR_ARG -1        /* invalid argument or malformed input       */
#define GCI_ERR_DEVICE -2     /* HIP runtime / device failure               */
#define GCI_ERR_NOMEM -3      /* device allocation failed or budget exceeded */
#define GCI_ERR_NOTFOUND -4   /* R^k integral missing (RKINTC would STOP)    */
#define GCI_ERR_NOCONV -5     /* Davidson hit MAXITER (DVDSON IERR 2048)     */
#define GCI_ERR_BUSY -6       /* serving: request not finished               */

typedef struct gci_hmat gci_hmat;       /* device-resident block Hamiltonian */
typedef struct gci_server gci_server;   /* continuous-batching solver       */

/* Last error message of the calling thread. */
const char *gci_last_error(void);

/* Select the device (hipSetDevice) and warm up the runtime. */
int gci_init(int device);

/* Page-locked host memory (hipHostMalloc / hipHostFree).  No counterpart in
 * the reference: result buffers for callers that keep eigenvectors in
 * pinned memory, which the device then writes by DMA with no staging copy. */
int gci_host_alloc(size_t bytes, void **ptr);
int gci_host_free(void *ptr);

/* ------------------------------------------------------------------ grid
 * Radial grid r_j = RNT (exp((j-1) H) - 1), RP = dr/dt, RPOR = RP / r,
 * n + 10 points (lib9290/radgrd.f90:41-60; rci90/getcid.f90:233-235). */
int gci_radgrd(int n, double rnt, double h, double *r, double *rp, double *rpor);

typedef struct {
    int n;              /* grid points N (590)                         */
    double h;           /* step H                                       */
    double accy;        /* H**6 (rci90/getcid.f90:262)                  */
    const double *r;    /* n + 10 */
    const double *rp;   /* n + 10 */
    const double *rpor; /* n + 10 */
} gci_grid;

typedef struct {
    int nw;             /* number of orbitals NW                        */
    int ld;             /* leading dimension of P, Q (>= n)             */
    const double *P;    /* nw x ld, orbital-major: P[j*ld + i] = PF(i+1,j+1) */
    const double *Q;
    const int32_t *mf;  /* MF: last tabulated point, 1-based            */
    const int32_t *nak; /* kappa                                        */
    const int32_t *nkl; /* l                                            */
    const int32_t *nkj; /* 2j                                           */
} gci_orbitals;

/* --------------------------------------------- piece 1: R^k tables
 * Replaces GENINTRK (rci90/genintrk.f90:3-109) and its SLATER/YZK/QUAD calls
 * (lib9290/slater.f90, yzk.f90, quad.f90).  Same enumeration order, key
 * packing ((ia*KEY+ib)*KEY+ic)*KEY+id with KEY = NW+1, and KSTART(0..j2max+1)
 * (kstart gets j2max+2 entries, 1-based positions).  Pass keys = vals = NULL
 * to get the count only. */
int gci_rk_table(const gci_grid *grid, const gci_orbitals *orb, int32_t *j2max,
                 int32_t *kstart, int32_t *keys, double *vals, int64_t cap, int64_t *n);

/* Hartree functions for a batch of (k, b, d) (lib9290/yzk.f90:3-223):
 * yk[q*n + i] = TB(i+1) and mtp[q] = MTP after YZK(k[q], b[q], d[q]),
 * orbitals 1-based.  Bit-identical to YZK (same marches, order and
 * association).  The reference's SCF program calls YZK one pair at a time
 * for its direct and exchange potentials (rmcdhf90/ypot.f90:75,
 * xpot.f90:80); this is that seam for a whole iteration's pairs. */
int gci_yzk(const gci_grid *grid, const gci_orbitals *orb, int npairs, const int32_t *k, const int32_t *b,
            const int32_t *d, double *yk, int32_t *mtp);

/* One-body Dirac integrals I(a,b) = RINTI(a,b,0) (lib9290/rinti.f90:49-103)
 * for every same-kappa pair a <= b, as IABINT memoises them
 * (rci90/iabint.f90).  zz: -r V_nuc(r), n entries (lib9290/nucpot.f90).
 * iab: nw x nw, iab[(a-1)*nw + (b-1)] for a <= b; other entries zero. */
int gci_one_body(const gci_grid *grid, const gci_orbitals *orb, const double *zz, double c,
                 double *iab);

/* --------------------------------------------- piece 2: H assembly
 * Angular lists exactly as the CPU producer emits them per CSF pair (ic, ir),
 * ir <= ic, pairs sorted by row then column (1-based):
 *   one-body  (ONESCALAR, librang90/onescalar.f90): IA, IB and TSHELL values
 *             at tsh_off (NW values if IA == IB, one if IA != IB, none if 0);
 *   two-body  (RKCO_GG/SPEAK, lib9290/speak.f90:46-59): entries
 *             [tb_beg, tb_end) of LABEL (a,b,c,d,k per entry) and COEFF.
 * Every row must carry its diagonal pair. */
typedef struct {
    int64_t npairs;
    const int32_t *pair_ic;
    const int32_t *pair_ir;
    const int32_t *one_ia;
    const int32_t *one_ib;
    const int64_t *tsh_off;
    const double *tshell;
    int64_t ntshell;
    const int64_t *tb_beg;
    const int64_t *tb_end;
    const int32_t *label;   /* 5 per two-body entry */
    const double *coeff;
    int64_t nterms;
} gci_pair_lists;

/* Replaces the SETHAM element loop (rci90/setham_gg.f90:162-385) for the
 * Dirac-Coulomb operator: per pair, ELEMNT = sum of T*I(a,b) over one-body
 * coefficients with |T| > 1e-12 plus V*R^k over two-body coefficients with
 * |V| > 1e-12 (R^k looked up as RKINTC does, rci90/rkintc.f90), summed in
 * list order; kept if diagonal or |ELEMNT| > 1e-12 (setham_gg.f90:364).
 * LFORDR/ICCUT: off-diagonal pairs with ir > iccut are skipped
 * (setham_gg.f90:172-174).  Produces the block Hamiltonian on the device
 * (diagonal shifted by -EAV as MANEIG loads it, maneig.f90:326) and returns
 * EAV = sum of diagonals / NCF (rci90/genmat2.f90:73) and the number of
 * stored lower-triangle elements (NELMNT). */
int gci_h_assemble(int ncf, int nw, const gci_pair_lists *lists, const double *iab,
                   const int32_t *kstart, const int32_t *keys, const double *vals,
                   int64_t nkeys, int lfordr, int iccut, int precision, gci_hmat **out,
                   double *eav, int64_t *nelmnt);

/* --------------------------------------------- transverse Breit
 * Replaces rci90's Breit radial integrals: GENINTBREIT1/2
 * (rci90/genintbreit1.f90:3-108, genintbreit2.f90; called at rci92.f90:198-
 * 202) and the memoised BRINT3..6 (rci90/brint3.f90..brint6.f90) with their
 * BRINTF/BESSEL/BRRA/RKINT/SKINT/ZKF chain.  The handle holds what the
 * reference keeps in module state across blocks: both tables, BESSEL's four
 * cached slots and frequencies (stor_C, bess_C; replayed in the reference's
 * call order, so every integral reads the slots the reference's would), and
 * the BRINT3..6 values.  e: orbital energies E (orb_C); c: speed of light;
 * wfact: WFACT (getcid.f90:154-164, photon frequency w = WFACT |E(a) -
 * E(b)| / c).  GENINTBREIT2's k = 0 entries go through ZKF(K-1 = -1), where
 * the reference reads an unset power table (zkf.f90:46-48); here
 * R**K = R**(-1) is used, and the angular code never asks for them. */
typedef struct gci_breit gci_breit;
int gci_breit_create(const gci_grid *grid, const gci_orbitals *orb, const double *e, double c, double wfact,
                     gci_breit **out);
/* The table of type 1 or 2 (keys ((ia*KEY+ib)*KEY+ic)*KEY+id, KEY = NW+1,
 * values, KSTARTBREIT(0..j2max+1)); keys = vals = NULL: count only. */
int gci_breit_table(const gci_breit *b, int type, int32_t *j2max, int32_t *kstart, int32_t *keys, double *vals,
                    int64_t cap, int64_t *n);
/* BRINTn(ia, ib, ic, id, k) as SETHAM calls it (1-based labels). */
int gci_breit_integral(gci_breit *b, int type, int ia, int ib, int ic, int id, int k, double *out);
int gci_breit_free(gci_breit *b);

/* SETHAM's transverse Breit lists (RKCO_GG with BREID, setham_gg.f90:298-
 * 302) per CSF pair of a gci_pair_lists: entries [bt_beg, bt_end) of LABEL
 * (6 per entry: a, b, c, d, k and the integral type 1..6, negative for a
 * core term) and COEFF. */
typedef struct {
    const int64_t *bt_beg;
    const int64_t *bt_end;
    const int32_t *label;
    const double *coeff;
    int64_t nterms;
} gci_breit_lists;

/* gci_h_assemble with the transverse Breit terms (setham_gg.f90:298-357):
 * after a pair's Dirac-Coulomb terms, COEFF * BRINTn over its Breit entries
 * with |COEFF| > 1e-12, in list order; core terms go to ELSTO instead of
 * the element.  EAV = sum of diagonals / NCF + ELSTO (genmat2.f90:73), and
 * the diagonal is shifted by -EAV (maneig.f90:326).  br carries the tables
 * and the state from block to block, as the reference's modules do: assemble
 * a run's blocks in the reference's block order. */
int gci_h_assemble_tb(int ncf, int nw, const gci_pair_lists *lists, const gci_breit_lists *breit,
                      gci_breit *br, const double *iab, const int32_t *kstart, const int32_t *keys,
                      const double *vals, int64_t nkeys, int lfordr, int iccut, int precision,
                      gci_hmat **out, double *eav, double *elsto, int64_t *nelmnt);

/* Streamed assembly: the same element loop (with or without Breit) fed in
 * chunks while the CPU producer makes the next ones.  The reference's SETHAM
 * interleaves production and summation row by row (setham_gg.f90:162-395);
 * here each gci_hstream_push copies one chunk of pairs (consecutive, in the
 * order of gci_pair_lists; offsets relative to the chunk's own arrays) into
 * a page-locked staging slot, queues its upload and the element kernel, and
 * returns: the copy and the kernel run while the caller produces the next
 * chunk.  Two slots alternate.  gci_hstream_end compacts the block; the
 * result is identical to gci_h_assemble(_tb) on the whole list for any
 * chunking.  br (may be NULL) makes it the Breit form; then every push
 * passes the chunk's Breit lists. */
typedef struct gci_hstream gci_hstream;
int gci_hstream_begin(int ncf, int nw, const double *iab, const int32_t *kstart, const int32_t *keys,
                      const double *vals, int64_t nkeys, int lfordr, int iccut, int precision, gci_breit *br,
                      gci_hstream **out);
int gci_hstream_push(gci_hstream *s, const gci_pair_lists *chunk, const gci_breit_lists *breit_chunk);
int gci_hstream_end(gci_hstream *s, gci_hmat **out, double *eav, double *elsto, int64_t *nelmnt);
int gci_hstream_free(gci_hstream *s);

/* --------------------------------------------- block Hamiltonian handle
 * .res-row ingestion (rci90/setham_gg.f90:378 layout): lower-triangle rows,
 * diagonal last, cumulative row ends iendc[0..n] (iendc[0] = 0), column
 * indices irow with the given index_base (1 for Fortran IROW).  The diagonal
 * is shifted by `shift` (MANEIG passes -EAV, maneig.f90:326).
 * precision: GCI_FP64 keeps FP64 off-diagonals (plain drop-in);
 * GCI_SPLIT32 also keeps hi = fp32(v) and lo = fp32(v - hi) for the mixed
 * form.  The diagonal is always FP64. */
#define GCI_FP64 0
#define GCI_SPLIT32 1
int gci_hmat_upload(int n, const int64_t *iendc, const int32_t *irow, const double *emt,
                    int index_base, double shift, int precision, gci_hmat **out);
int gci_hmat_free(gci_hmat *h);
int gci_hmat_info(const gci_hmat *h, int64_t *n, int64_t *nnz_offdiag, int *precision);
/* Download as lower rows with diagonal last (inverse of the upload). */
int gci_hmat_download(const gci_hmat *h, int64_t *iendc, int32_t *irow, double *emt,
                      int index_base);

/* --------------------------------------------- piece 3: SPICMV2
 * C := H B for the symmetric H held as lower rows (lib9290/spicmv2.f90:2-60,
 * dmerge.f90:36-40): B, C column-major n x m, C overwritten.  This is the
 * GDVD operator contract OP(N, M, B, C) (libdvd90/gdvd.f90:139).
 * mode: GCI_OP_FP64 (FP64 values), GCI_OP_SPLIT_BULK (fp32 hi values and
 * fp32-rounded B, FP64 accumulation, FP64 diagonal), GCI_OP_SPLIT_FULL
 * (hi + lo, FP64 B). */
#define GCI_OP_FP64 0
#define GCI_OP_SPLIT_BULK 1
#define GCI_OP_SPLIT_FULL 2
int gci_spmm(gci_hmat *h, int m, const double *B, double *C, int mode);
/* Same on device pointers (leading dimension ld >= n), enqueued on the
 * handle's stream; no synchronisation.  The handle's stream is
 * non-blocking: work that produced dB on another stream must have completed
 * (or be waited for) before the call. */
int gci_spmm_device(gci_hmat *h, int m, const double *dB, int ldb, double *dC, int ldc, int mode);

/* No counterpart in the reference (SPICMV2 allocates nothing, and the GDVD
 * contract has no workspace for the operator): the same operator as
 * gci_spmm_device for callers that cannot allocate device memory -- work
 * captured into a HIP graph, tight serving loops, harnesses that forbid
 * hipMalloc.  It calls no hipMalloc/hipFree: its staging (n x launch width
 * rows of x and y, launch width <= 8) and its launch descriptor live in one
 * library-owned array of GCI_NOALLOC_SCRATCH_BYTES (48 MiB) per device, so
 * n x width may be at most about 3.1e6 (n <= 393k rows at 8 columns);
 * larger calls return GCI_ERR_ARG.  The matrix must be prepared
 * beforehand: gci_spmm_layout (or one gci_spmm_device call) for the launch
 * shape in force, and for GCI_OP_FP64 on a split matrix one gci_spmm_device
 * call in that mode.  Enqueued on the handle's stream, no synchronisation.
 * Every call on a device shares that device's array, which keeps the last
 * call's contents: calls that overlap on one device race. */
#define GCI_NOALLOC_SCRATCH_BYTES (48u << 20)
int gci_spmm_device_noalloc(gci_hmat *h, int m, const double *dB, int ldb, double *dC, int ldc, int mode);

/* --------------------------------------------- piece 4: Davidson
 * GDVD/DVDSON contract (libdvd90/gdvd.f90:4-147, dvdson.f90): lowest nvec
 * eigenpairs of H (as held, i.e. already shifted), LIM, CRITE, CRITC, CRITR,
 * MAXITER, MBLOCK as in MANEIG (maneig.f90:389-420).  Gram-Schmidt, the
 * Rayleigh-Ritz projection and basis rotations run on the device. */
typedef struct {
    int nvec;           /* IHIGH (ILOW = 1)                                 */
    int lim;            /* LIM; 0 = min(n, 2*nvec+60)                       */
    int mblock;         /* MBLOCK; 0 = 1 (in-core rci)                       */
    int maxiter;        /* MAXITER; 0 = max(200*nvec, n/10)                  */
    double crite;       /* 0 = 1e-17                                         */
    double critc;       /* 0 = 1e-9                                          */
    double critr;       /* 0 = 1e-9                                          */
    int niv;            /* initial vectors supplied in evec (0 = INIEST2)     */
    int mode;           /* operator mode (GCI_OP_*)                          */
    int refine;         /* split forms: FP64 refinement passes at the end    */
} gci_dvd_params;

typedef struct {
    int nloops;         /* NLOOPS */
    int nmv;            /* NMV: matrix-vector products (columns)             */
    int ierr;           /* DVDSON IERR bitmask                               */
    int refine_loops;   /* FP64 iterations after the split-precision phase   */
    double seconds;     /* wall time of the solve                            */
} gci_dvd_stats;

/* eval[nvec] (eigenvalues of H as held), evec n x nvec column-major,
 * resid[nvec] residual norms (may be NULL). */
int gci_davidson(gci_hmat *h, const gci_dvd_params *p, double *eval, double *evec,
                 double *resid, gci_dvd_stats *st);

/* MANEIG dispatch for one block (rci90/maneig.f90:130-555): NCF == 1
 * trivial, NCF <= 2000 dense eigensolve, else INIEST2 + GDVD(SPICMV2);
 * eigenvectors sign-fixed so the largest component is positive
 * (maneig.f90:532-555). */
int gci_maneig(gci_hmat *h, int nvec, double *eval, double *evec, gci_dvd_stats *st);

/* MANEIG of the SCF program rmcdhf90 (rmcdhf90/maneig.f90:78-200), the
 * second caller of GDVD(SPICMV2) (:161).  Computes the lowest
 * NVEX = min(nvect, NCF) eigenpairs (nvect = NCMAXBLK, at most 16) and
 * returns the levels iccmin[0..nlev) (1-based, ICCMIN): eval[nlev],
 * evec n x nlev, sign-fixed (:179-196).  Cold start (warm == 0, the first
 * SCF iteration) or NCF <= 4000: INIEST2 over the leading min(4000, NCF)
 * block, which for NCF <= 4000 is the answer.  Warm start: eval/evec hold
 * the previous iteration's levels on entry (orthonormal vectors, as INITDVD
 * takes its start block unchanged) and start GDVD (:118-124); the reference
 * leaves WORK's other start columns as allocated, here column c is the unit
 * vector at the c-th smallest diagonal element, orthonormalised against the
 * given columns and the filled ones before it.  GDVD runs with
 * LIM = min(NCF, 2*nvect+40), MAXITR = min(200*nvect, NCF), MBLOCK = 1,
 * CRITE 1e-17, CRITC = CRITR = 1e-9 (:104-162). */
int gci_maneig_mcdhf(gci_hmat *h, int nvect, int nlev, const int32_t *iccmin, int warm, double *eval,
                     double *evec, gci_dvd_stats *st);

/* --------------------------------------------- batched + serving
 * Many blocks at once: each block's Davidson iterations are interleaved on
 * the device (split-precision bulk operator, FP64 refinement).  tol[b] is
 * the relative eigenvalue tolerance of block b; eval_total adds back
 * shift[b] (EAV) so errors are relative to total energies. */
int gci_batch_solve(int nblocks, gci_hmat **blocks, const int *nev, const double *tol,
                    const double *eav, double *eval, int64_t *eval_off, double *evec,
                    int64_t *evec_off, gci_dvd_stats *stats);

int gci_server_create(size_t hbm_budget_bytes, int max_active, gci_server **out);
int gci_server_destroy(gci_server *s);
/* Submit one CI solve; the server takes ownership of h.  The matrix's SpMM
 * layout is built in this call (on the caller's thread), so the latency of
 * a request runs from here.  Returns a request id in *req. */
int gci_server_submit(gci_server *s, gci_hmat *h, int nev, double tol, double eav, int64_t *req);
/* The same for a matrix the caller keeps: the server does not free h, several
 * requests may share it, and only their Davidson workspace counts against
 * the HBM budget.  h must outlive every request submitted with it. */
int gci_server_submit_shared(gci_server *s, gci_hmat *h, int nev, double tol, double eav, int64_t *req);
/* Non-blocking poll: GCI_OK with results when done, GCI_ERR_BUSY otherwise. */
int gci_server_poll(gci_server *s, int64_t req, double *eval, double *evec, double *latency_s);
/* Block until every submitted request has finished. */
int gci_server_drain(gci_server *s);

/* --------------------------------------------- operator profiling
 * No counterpart in the reference (which only times whole programs with
 * STARTTIME/STOPTIME): per operator mode (GCI_OP_*), the number of symmetric
 * SpMM kernel launches made by the Davidson engines, their summed kernel
 * time in milliseconds (HIP events on the launching stream) and the summed
 * algorithmic bytes of the blocks they served.  While enabled, engine
 * iterations are launched directly instead of from a captured graph.
 * stats: 9 doubles, [launches, ms, bytes] x 3 modes. */
int gci_profile_enable(int on);
int gci_profile_read(double *stats);

/* --------------------------------------------- SpMM launch shape, autotuner
 * No counterpart in the reference (SPICMV2 has no tunables).  The symmetric
 * SpMM's launch shape: panel height R (1024, 2048 or 4096: rows whose x
 * (and, for lower-triangle layouts, y) stay in LDS; the panel-slice layout
 * is built for it), workgroup size (256, 512, 1024), at most `nslice` slices
 * (64-column groups, whose column tables are staged in LDS) per work item of
 * a full layout (at most 128) and per LDS chunk of a lower-triangle one,
 * about `target` entries per work item (whole slices, or a run of steps, of
 * one panel), and steps each wavefront keeps in flight `depth` (2, 4, 8).
 * One shape is in force per process; a matrix laid out for another R is laid
 * out again at its next use outside a Davidson engine.  Set or tune the shape
 * before solves, not while requests are resident (an engine refuses a change
 * of R under resident problems). */
typedef struct gci_spmm_shape {
    int R, block, nslice, depth;
    int64_t target;
} gci_spmm_shape;
int gci_spmm_get_shape(gci_spmm_shape *out);
int gci_spmm_set_shape(const gci_spmm_shape *shape);
/* The panel-slice layout of a matrix for the shape in force (laid out now if
 * it was not): its column segments, slices and stored entries (pads and
 * escape entries included: stored / laid-out entries is the stream's
 * overhead). */
int gci_spmm_layout(gci_hmat *h, int64_t *nseg, int64_t *nslice, int64_t *stored);
/* Mean milliseconds of the SpMM kernel launches of one m-column operator
 * application (mode GCI_OP_*) to every matrix of the set at once -- all
 * matrices' work items in one launch per column block, as the batched
 * Davidson engine issues them -- on seeded device-resident operands; `reps`
 * timed repetitions after one untimed application. */
int gci_spmm_time(gci_hmat *const *mats, int nmats, int m, int mode, int reps, double *ms);
/* Coordinate search over the shape parameters on that timing, starting from
 * the shape in force; leaves the fastest shape in force and returns it and
 * its time. */
int gci_spmm_autotune(gci_hmat *const *mats, int nmats, int m, int mode, int reps, gci_spmm_shape *best,
                      double *best_ms);

#ifdef __cplusplus
}
#endif
#endif
