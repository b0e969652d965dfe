// dvd.h -- device-resident block Davidson (GDVD/DVDSON) over many problems.
#pragma once
#include "gci_internal.h"

#include <algorithm>

namespace gci {

constexpr int MAXNUME = 16;   // largest number of wanted eigenpairs
// largest LIM: MANEIG's 2*NVECMX+60 for every NVECMX <= MAXNUME
// (rci90/maneig.f90:389), rounded up to whole 16-column MFMA tiles
constexpr int MAXLIM = 96;
static_assert(MAXLIM >= 2 * MAXNUME + 60 && MAXLIM % 16 == 0, "MAXLIM must hold 2*MAXNUME+60 in 16-column tiles");
// NINCOR of MANEIG (rci90/maneig.f90:93)
constexpr int64_t NINCOR = 268435456;
constexpr int CHUNK = 512;    // rows per tall-skinny work item
constexpr int RESULT_STRIDE = 5 + 2 * MAXNUME;

// Per-problem stage.  Kernels act only on problems in the stage they serve.
enum : int {
    PH_EMPTY = -1,     // free slot
    PH_ITER = 0,       // inside DVDRVR's loop
    PH_FINAL = 1,      // converged / stopped: exit rotation + residuals pending
    PH_DONE = 2,       // results ready
    PH_INIT = 3,       // INITDVD step in progress
    PH_INIT_IDLE = 4,  // INITDVD, nothing to do in the current sub-step
};

// One Davidson problem (a CI block).  Lives in device memory; the scalar
// state is advanced by device kernels so an iteration needs no host sync.
struct DvdProb {
    // --- static description
    int64_t n;
    int lim, nume, mblock, maxiter;
    double crite, critc, critr;
    int mode;                       // SpMM operator mode (GCI_OP_*) of the bulk phase
    const int64_t *rp;              // off-diagonal CSR of H
    const int32_t *col;
    const double *val;
    const float *hi, *lo;
    const double *diagH;            // FP64 diagonal of H (shifted)
    const int32_t *pscol;           // panel-slice view used by the operator
    const int64_t *psstep;
    const uint16_t *plrow;
    const double *pval;
    const float *phi, *plo;
    int64_t pnnz;                   // stored off-diagonal entries (lower triangle)
    int pfull;                      // the layout holds both triangles
    const uint8_t *pdrow;           // full layouts: row deltas and first rows
    const uint16_t *psfirst;
    void *xs;                       // operator staging: n x 8 (fp32 or fp64)
    double *ys;                     // n x 8 fp64
    double *basis, *ab;             // n x lim each (column-major, ld = n)
    double *diag;                   // preconditioner diagonal (drifts as in GDVD)
    double *xc, *rc;                // n x nume candidate Ritz vectors / residuals
    double tol_rel;                 // >0: residual test ||r|| < tol_rel*|lambda+eshift|
    double eshift;                  // EAV added back for the relative test
    int64_t chunk0;                 // first entry in the flattened chunk list
    int nchunks;
    int slot;
    double tol_final, critr_final;  // targets of the FP64 refinement pass
    // --- dynamic state
    int kpass, nncv, nnew, rest, nloops, ierr, nmv;
    int phase, ovflow, ncand, conv_all, niv;
    int refine_phase;               // 0 bulk operator, 1 FP64 (hi+lo) operator
    int will_refine;                // an FP64 refinement follows the bulk phase
    int refine_loops;
    int bulk_loops, bulk_nmv;       // iterations/products of the bulk phase
    int brk_mask;                   // new columns that lost rank in the first GS pass
    int skip2;                      // the first GS pass kept every column's norm (DGKS): no second pass
    int init_ist, init_iend;        // INITDVD progress: next AB column to copy, columns built
    int init_src;                   // AB column copied by the current INITDVD sub-step
    // --- small arrays (1-based reference arrays stored 0-based)
    double eigval[MAXLIM];          // all KPASS Ritz values, ascending
    double oldval[MAXLIM + 1];
    double svec[MAXLIM * MAXLIM];   // kpass x kpass Ritz coefficients, ld = kpass
    double s[MAXLIM * (MAXLIM + 1) / 2];
    double resn[MAXLIM];            // SCRA1(1:NUME) residual norms
    double epsil[MAXLIM];           // SCRA1(LIM+1:) eigenvalue shifts (zero)
    int icv[MAXNUME + 2];
    int incv[MAXLIM + 1];
    int slot_cand[MAXNUME];
    double rhseps[MAXNUME];
    double lam_add[MAXNUME];
    double coef[MAXLIM * MAXNUME];
    double rinv[MAXNUME * MAXNUME];
    int unit_row[MAXNUME];          // rows of the unit start vectors (k_unit_merge)
};

// Unit start vectors of a batched or served solve.  NUME + 1 of them replace
// INITDVD's expansion step (orthogonalising A b_1 and one more operator
// product) when they still fit one 8-wide operator launch; past that the
// Krylov direction A b_1 is the better extra vector (an 8-pair block took
// 13 bulk iterations from 9 unit vectors against 11 with the expansion).
inline int start_vectors(int nume, int lim) { return nume + 1 <= 8 ? std::min(nume + 1, lim) : nume; }

// Residual target of the split-precision bulk phase for a final target
// tol.  The 2e-10 floor is below the fp32 rounding of H on purpose: a bulk
// phase stopped at 1e-8 or 3e-7 left the FP64 refinement more iterations and
// the batched solve slower (22.8 and 23.7 ms against 21.1 ms).
inline double bulk_tolerance(double tol) { return std::max(0.5 * tol, 2.0e-10); }

struct DvdSpec {
    const HMat *h = nullptr;
    int nume = 1, lim = 0, mblock = 1, maxiter = 0;
    double crite = 1e-17, critc = 1e-9, critr = 1e-9;
    int mode = GCI_OP_FP64;
    double tol_rel = 0.0;   // bulk-phase relative residual target (0: CRITR)
    double tol_final = 0.0; // refinement target (split forms)
    double critr_final = 0.0;
    double eshift = 0.0;
    bool refine = false;    // run an FP64 refinement after the bulk phase
};

// Davidson over a fixed-capacity set of problem slots sharing one stream.
// Problems can be admitted and released between steps (continuous
// batching); an iteration is one fixed launch sequence over all slots,
// replayed from a captured HIP graph while the slot set is unchanged.
class DavidsonEngine {
  public:
    DavidsonEngine(int capacity, hipStream_t s);
    ~DavidsonEngine();
    DavidsonEngine(const DavidsonEngine &) = delete;
    DavidsonEngine &operator=(const DavidsonEngine &) = delete;

    // Admits problems; init[i] = host n x niv[i] vectors, or null for unit
    // vectors on the niv smallest diagonal entries.  Runs INITDVD.
    std::vector<int> admit(const std::vector<DvdSpec> &specs, const std::vector<const double *> &init,
                           const std::vector<int> &niv);
    void step(int iters);                 // enqueue iterations (no sync)
    std::vector<int> poll();              // sync; slots newly in PH_DONE
    // Slots newly in PH_DONE as of the iteration before the last step():
    // waits only for that one, so the device keeps the last step's work
    // queued while the host decides (every iteration ends by writing the
    // slots' phases to mapped host memory).
    std::vector<int> poll_lagged();
    // every admitted slot is done and needs no refinement pass
    bool all_done() const;
    // Restart finished bulk-phase slots with the FP64 operator from their
    // Ritz vectors (the refinement pass).
    void refine_restart(const std::vector<int> &slots);
    bool needs_refine(int slot) const;
    void extract(int slot, double *eval, double *evec, double *resid, gci_dvd_stats *st);
    void release(int slot);
    void release_all();
    // compact per-slot results: [nloops, nmv, ierr, refined, bulk_loops,
    // eigval(16), resn(16)] (RESULT_STRIDE doubles per slot)
    void extract_all(std::vector<double> &res);
    // eigenvectors of the given slots, one after another (n x NUME each,
    // vector-major), into evec; returns when the copy has completed
    void copy_vectors(const std::vector<int> &slots, double *evec);
    void sync() {
        GCI_HIP(hipStreamSynchronize(stream_));
        waited_seq_ = launched_;
        drain_profile(true);
    }
    int active() const;
    // admitted slots still iterating (neither done nor empty), as of the last poll
    int iterating() const;
    int capacity() const { return cap_; }
    int phase(int slot) const { return phase_[slot]; }
    static size_t workspace_bytes(const DvdSpec &s);
    // admission rules of one problem started from niv vectors (GCI_ERR_ARG)
    static void check_spec(const DvdSpec &s, int niv);
    bool use_graph = true;

  private:
    void rebuild_chunks();
    void launch_iteration();
    void launch_finalize();
    void launch_orth(int passes, int phase);
    void launch_gram(int G, int sel, int phase, int pass2);
    void launch_op(int phase, int M, int modes = -1);  // modes: mask of operator modes (-1: all in use)
    void update_modes();
    // unit_op: every starting problem has unit start vectors (k_unit_merge),
    // whose first product is read from H's columns instead of an SpMM pass
    void run_initdvd(const std::vector<int> &slots, bool unit_op = false);
    void drain_profile(bool wait);
    hipEvent_t pooled_event();
    struct ProfEv {
        int mode;
        hipEvent_t e0, e1;
    };
    std::vector<ProfEv> prof_events_;
    std::vector<hipEvent_t> ev_pool_;  // events of drained profile pairs, reused

    int cap_;
    hipStream_t stream_;
    std::vector<DvdSpec> spec_;
    std::vector<int> phase_;               // last polled phases
    std::vector<char> refined_;
    std::vector<int> niv_;
    std::vector<DBuf<double>> work_;
    DBuf<DvdProb> dprob_;
    DBuf<int> chunk_prob_;
    DBuf<int64_t> chunk_row_;
    DBuf<double> partial_;
    DBuf<int> dphase_;
    DBuf<SpmmJob> djobs_;
    DBuf<SpmmItem> ditems_;
    DBuf<char> dmeta_;  // per-slot chunk ranges, staged for k_set_chunks
    // Persistent staging, so admissions and result reads allocate nothing
    // (a temporary's hipFree synchronises the device) and copy from pinned
    // memory (pageable copies block the host one by one).
    char *hstage_pin_ = nullptr;    // admitted problems' initial state (DvdProb heads)
    int *hwant_pin_ = nullptr;      // unit-vector start flags
    double *hres_pin_ = nullptr;    // gathered results
    void *hrec_pin_ = nullptr;      // admitted problems' slots and workspace sizes
    int *hphase_pin_ = nullptr;     // phases read by poll()
    // iteration snapshot: [sequence number, phase of each slot], written by
    // the last kernel of every iteration into mapped host memory
    int *hsnap_pin_ = nullptr;
    int *dsnap_host_ = nullptr;     // device address of hsnap_pin_
    DBuf<int> dseq_;                // iterations completed (device counter)
    int64_t launched_ = 0;          // iterations launched
    std::vector<int64_t> restart_seq_;  // snapshots up to this one predate a slot's (re)start
    static constexpr int NEV = 4;
    hipEvent_t ev_[NEV] = {};
    int64_t ev_seq_[NEV] = {};      // iterations launched when ev_[i] was recorded
    int nev_rec_ = 0;               // events recorded
    int *hmask_pin_ = nullptr;      // refinement restart flags (NEV rows of cap_), mapped
    int *dmask_host_ = nullptr;
    int mask_row_ = 0;
    int64_t mask_seq_[NEV] = {-1, -1, -1, -1};  // launched_ when each flag row was last written
    int64_t waited_seq_ = 0;        // iterations known complete
    char *hchunk_pin_ = nullptr;    // chunk and item tables staged for upload
    size_t hchunk_cap_ = 0;
    hipEvent_t chunk_ev_ = nullptr; // their last upload
    bool chunk_ev_rec_ = false;
    hipEvent_t stage_ev_ = nullptr;  // the last admission's staging copies (pinned staging reuse)
    bool stage_ev_rec_ = false;
    double *hvec_pin_ = nullptr;    // staging of copy_vectors()
    size_t hvec_bytes_ = 0;
    std::vector<hipEvent_t> vec_ev_;
    int64_t *hvoff_pin_ = nullptr;  // per-slot offsets for copy_vectors()
    DBuf<int64_t> dvoff_;
    DBuf<double> dvec_;             // gathered eigenvectors
    DBuf<int> dwant_;
    DBuf<char> dstage_, drec_;      // admitted problems, staged contiguously
    DBuf<double> dres_;
    int64_t nitems_ = 0;
    int modes_used_ = 0;           // bitmask of operator modes over active slots
    int mwidth_iter_ = 1;          // launch width for iterations
    int nume_max_ = MAXNUME;       // largest NUME over active slots (k_newvec's accumulators)
    int lim_max_ = MAXLIM;         // largest LIM over active slots (k_gram's tile width)
    int64_t total_chunks_ = 0;
    std::vector<int> hprob_;
    std::vector<int64_t> hrow_;
    hipGraphExec_t gexec_ = nullptr;
    bool graph_valid_ = false;
    uint64_t shape_version_ = 0;  // SpMM shape the captured graph was built for
    // the captured iteration carries OVFLOW's rotation and reset (k_rotate,
    // k_ovf_fix): some slot's KPASS may reach LIM in the iterations queued
    bool ovf_in_graph_ = true;
    bool ovflow_possible(int iters) const;
};

// Operator profiling (gci_profile_*): per operator mode, launches, summed
// kernel milliseconds (HIP events around the panel-slice kernel) and summed
// algorithmic bytes of the jobs each launch served.
struct OpProfile {
    bool on = false;
    double launches[3] = {0, 0, 0};
    double ms[3] = {0, 0, 0};
    double bytes[3] = {0, 0, 0};
};
OpProfile &op_profile();
double *op_bytes_device();  // device accumulator [3] written by k_make_jobs

// Dense lowest eigenpairs on the device (rocSOLVER dsyevx): a is n x n
// column-major (device), eval/evec device outputs for il..iu (1-based).
void dense_lowest(int n, double *d_a, int nev, double *d_eval, double *d_evec, hipStream_t s);

}  // namespace gci
