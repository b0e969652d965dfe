/*
 * graspci_oracle.h -- CPU restatement of GRASP2018's rci90 hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * engine in grasp_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path never links
 * or calls it.
 *
 * Every function names the reference routine (file:line, relative to the
 * GRASP source tree) whose arithmetic it restates.  Array conventions:
 * C arrays are 0-based; element t holds the Fortran element t+1.  Orbital
 * tables are orbital-major: P[j*ld + i] is PF(i+1, j+1).  Orbital labels in
 * angular lists stay 1-based, exactly as the Fortran producer emits them.
 */
#ifndef GRASPCI_ORACLE_H
#define GRASPCI_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* radial grid, exponential branch of RADGRD (lib9290/radgrd.f90:41-60);
 * r, rp, rpor have n+10 entries. */
void orc_radgrd(int n, double rnt, double h, double *r, double *rp, double *rpor);

/* Newton-Cotes constants of SETQIC (lib9290/setqic.f90:312-322).
 * ncc[0..3] = C1..C4; cnc5c[(j-2)*4 + (i-2)] = CNC5C(i,j), i=2..5, j=2..4. */
void orc_setqic(double h, double *ncc, double *cnc5c);

/* 13-point Lagrange derivative coefficients A(13,13) of SETQIC
 * (setqic.f90:179-234, 291-296); a13[(i-1)*13 + (j-1)] = A(i,j). */
void orc_lagrange13(double *a13);

/* QUAD (lib9290/quad.f90:3-110).  ta has at least mtp+3 entries and is
 * modified (tail zeroed) as in the reference. */
double orc_quad(const double *r, const double *rp, const double *ncc,
                double *ta, int mtp);

/* YZK (lib9290/yzk.f90:3-226): Hartree Y^k of orbitals i, j (0-based
 * indices into P/Q, each row ld long).  yk gets n entries; returns MTP. */
int orc_yzk(int n, const double *r, const double *rp, const double *ncc,
            const double *cnc5c, double accy, int k,
            const double *P, const double *Q, int ld, const int *mf,
            int i, int j, double *yk);

/* SLATER (lib9290/slater.f90:3-81): R^k(ia,ib;ic,id), 1-based labels. */
double orc_slater(int n, const double *r, const double *rp, const double *rpor,
                  const double *ncc, const double *cnc5c, double accy,
                  const double *P, const double *Q, int ld, const int *mf,
                  int ia, int ib, int ic, int id, int k);

/* ES + NUCPOT (lib9290/es.f90, lib9290/nucpot.f90): -r V(r) in zz[0..n-1].
 * nparm = 0 (point) or 2 (Fermi, parm = {c, a} in bohr). */
void orc_es(double f, double *s2f, double *s3f);
void orc_nucpot(int n, const double *r, double z, int nparm, const double *parm,
                double *zz);

/* DPBDT (lib9290/dpbdt.f90): ta/tb = d/dt P, Q of orbital j (0-based). */
void orc_dpbdt(int n, const double *a13, const double *P, const double *Q,
               int ld, int j, double *ta, double *tb);

/* RINTI (lib9290/rinti.f90:3-119): one-body Dirac integral I(j,k),
 * 1-based labels; nak = kappa per orbital; c = speed of light. */
double orc_rinti(int n, const double *r, const double *rp, const double *rpor,
                 const double *ncc, const double *a13, double h, double c,
                 const double *zz, const double *P, const double *Q, int ld,
                 const int *mf, const int *nak, int j, int k, int mode);

/* TRIANGRK (rci90/triangrk.f90) */
int orc_triangrk(int la, int k, int lb);

/* GENINTRK enumeration (rci90/genintrk.f90:40-86).  nkl = l, nkj = 2j per
 * orbital.  kstart gets KSTART(0..j2max+1) (1-based positions, as in the
 * reference).  keys may be NULL (count only).  Returns the number of R^k. */
int64_t orc_genintrk_enum(int nw, const int *nkl, const int *nkj, int *j2max,
                          int32_t *kstart, int32_t *keys, int32_t *kval,
                          int32_t *labels /* 4 per entry, 1-based */);

/* RKINTC (rci90/rkintc.f90:40-103).  Returns 0 and the value, or -1 when the
 * integral is not in the table (the reference STOPs). */
int orc_rkintc(int nw, const int32_t *kstart, const int32_t *keys,
               const double *vals, int ia, int ib, int ic, int id, int k,
               double *val);

/* SETHAM reduction (rci90/setham_gg.f90:162-395) over CPU-produced angular
 * lists.  Pairs are (ic, ir) with ir <= ic, rows ascending, columns ascending
 * within a row, 1-based.  One-body record per pair: ONESCALAR's IA, IB and
 * the TSHELL values starting at tsh_off (NW values when IA == IB, one when
 * IA != IB, none when IA == 0).  Two-body entries [tb_beg, tb_end): LABEL
 * (5 per entry: a, b, c, d, k) and COEFF.  iab[(a-1)*nw + (b-1)] = I(a,b).
 * Output: lower CSR with diagonal last (row_end[0] = 0, cols 1-based),
 * returns nnz or -1 on a missing integral / malformed list; *eav_sum gets
 * the diagonal sum (EAV before the division by NCF). */
int64_t orc_setham(int ncf, int nw, int64_t npairs,
                   const int32_t *pair_ic, const int32_t *pair_ir,
                   const int32_t *one_ia, const int32_t *one_ib,
                   const int64_t *tsh_off, const double *tshell,
                   const int64_t *tb_beg, const int64_t *tb_end,
                   const int32_t *label, const double *coeff,
                   const double *iab,
                   const int32_t *kstart, const int32_t *keys, const double *vals,
                   int lfordr, int iccut,
                   int64_t *row_end, int32_t *col, double *val, double *eav_sum);

/* ------------------------------------------------- transverse Breit
 * State of rci90's Breit radial integrals (breit_oracle.c): grid, orbitals,
 * orbital energies E (orb_C), C and WFACT (photon frequencies w = WFACT
 * |E(a) - E(b)| / C, bessel.f90:80), BESSEL's cache (KEEP, WIJ, BESSJ,
 * BESSN), the GENINTBREIT1/2 tables and the BRINT3..6 memo.  All arrays are
 * the caller's and must outlive the state.  Labels are 1-based. */
typedef struct orc_breit orc_breit;
orc_breit *orc_breit_new(int n, const double *r, const double *rp, const double *rpor, const double *ncc,
                         const double *cnc5c, double accy, int nw, int ld, const double *P, const double *Q,
                         const int *mf, const int *nak, const int *nkl, const int *nkj, const double *e,
                         double c, double wfact);
void orc_breit_free(orc_breit *s);
/* BRINTF (rci90/brintf.f90:44-96): one integral of type 1..6, through
 * BESSEL's cache (state-dependent as the reference is). */
double orc_brintf(orc_breit *s, int itype, int ia, int ib, int ic, int id, int k);
/* TRIANGBREIT1 (type 1) / TRIANGBREIT2 (type 2) on (l, 2j) of a, b, c, d. */
int orc_triangbreit(int type, int la, int ja, int lb, int jb, int lc, int jc, int ld, int jd, int k);
/* GENINTBREIT1 / GENINTBREIT2 (rci90/genintbreit1.f90:45-104): the table of
 * type 1 or 2 (keys ((ia*KEY+ib)*KEY+ic)*KEY+id, KEY = NW+1, values,
 * KSTARTBREIT(0..j2max+1)), kept in the state for BRINT1/2.  keys/vals may
 * be NULL.  Returns the count, -1 if J2MAX > KMAX (20). */
int64_t orc_genintbreit(orc_breit *s, int type, int *j2max, int *kstart, int32_t *keys, double *vals);
/* The table of type 1 or 2 held in the state; returns its length. */
int64_t orc_breit_table(const orc_breit *s, int type, int32_t *keys, double *vals);
/* BRINT1..6 (rci90/brint1.f90, brint3.f90): table lookup (1, 2) or the
 * memoised BRINTF (3..6).  Returns 0, -1 for a missing table integral. */
int orc_brint(orc_breit *s, int type, int ia, int ib, int ic, int id, int k, double *out);
/* SETHAM's Breit terms of one pair (setham_gg.f90:298-357): entries
 * [beg, end) of LABEL (6 per entry: a, b, c, d, k, type; type < 0: core)
 * and COEFF. */
int orc_setham_breit_pair(orc_breit *s, int64_t beg, int64_t end, const int32_t *label, const double *coeff,
                          double *elemnt, double *elsto);
/* orc_setham with the transverse Breit terms of every pair ([bt_beg,
 * bt_end) of blabel/bcoeff, summed after the Dirac-Coulomb terms) and the
 * core sum ELSTO (*elsto).  br == NULL: Dirac-Coulomb only. */
int64_t orc_setham_tb(int ncf, int nw, int64_t npairs,
                      const int32_t *pair_ic, const int32_t *pair_ir,
                      const int32_t *one_ia, const int32_t *one_ib,
                      const int64_t *tsh_off, const double *tshell,
                      const int64_t *tb_beg, const int64_t *tb_end,
                      const int32_t *label, const double *coeff,
                      const double *iab,
                      const int32_t *kstart, const int32_t *keys, const double *vals,
                      const int64_t *bt_beg, const int64_t *bt_end,
                      const int32_t *blabel, const double *bcoeff, orc_breit *br,
                      int lfordr, int iccut,
                      int64_t *row_end, int32_t *col, double *val, double *eav_sum,
                      double *elsto);

/* SPICMV2 + DMERGE (lib9290/spicmv2.f90:2-60, dmerge.f90:3-43).
 * row_end[0..n] cumulative ends (IENDC with IENDC(0)=0), cols 1-based,
 * diagonal last in each row.  B, C column-major n x m. */
void orc_spicmv2(int n, int m, const int64_t *row_end, const int32_t *col,
                 const double *val, const double *B, double *C);

/* Same operator, row-cyclic over nthreads threads with a private output per
 * thread and a reduction, mirroring mpi90/spicmvmpi.f90:61-75. */
void orc_spicmv2_mt(int n, int m, const int64_t *row_end, const int32_t *col,
                    const double *val, const double *B, double *C, int nthreads);

/* Dense symmetric eigensolver (Householder tridiagonalisation + implicit QL;
 * the published tred2/tql2 method of Wilkinson & Reinsch).  a is n x n
 * column-major, overwritten with eigenvectors; w ascending eigenvalues.
 * Stands in for LAPACK DSPEVX at its call sites (dvdson.f90:724,
 * maneig.f90:220, iniest2.f90).  Returns 0 on success. */
int orc_syev(int n, double *a, double *w);

/* Lowest nev eigenpairs by the steps of LAPACK's DSPEVX (RANGE='I', IL=1,
 * IU=nev; called at rci90/maneig.f90:220 and in lib9290/iniest2.f90):
 * tridiagonal reduction, bisection, inverse iteration, back-transformation.
 * a: n x n column-major, both triangles (destroyed); w[nev]; z n x nev. */
int orc_spevx_lowest(int n, double *a, int nev, double *w, double *z, int nthreads);

/* GDVD + DVDSON (libdvd90/gdvd.f90, libdvd90/dvdson.f90) for the lowest
 * ILOW=1..IHIGH eigenpairs with the CSR operator above.  work must hold
 * n*nume initial vectors (niv of them filled) on entry and is sized by
 * orc_gdvd_worksize.  On exit work[0..n*nume) = eigenvectors, then nume
 * eigenvalues, eigenvalue differences, residual norms (dvdson.f90:465-474).
 * diag is modified exactly as GDVD modifies it. */
int64_t orc_gdvd_worksize(int n, int lim, int nume);
int orc_gdvd(int n, const int64_t *row_end, const int32_t *col, const double *val,
             int lim, double *diag, int ihigh, int niv, int mblock,
             double crite, double critc, double critr, int maxiter,
             double *work, int nthreads, int *nloops, int *nmv);

/* INIEST2 (lib9290/iniest2.f90): lowest niv eigenvectors of the leading
 * min(nmax, ncf) block, zero padded, written to basis (ncf x niv). */
int orc_iniest2(int nmax, int ncf, int niv, const int64_t *row_end,
                const int32_t *col, const double *val, double *basis);

/* NINCOR of MANEIG (rci90/maneig.f90:93): matrices whose NSTORE exceeds it
 * are kept on disk, and GDVD then runs with MBLOCK = NVEX. */
#define ORC_NINCOR 268435456LL

/* MANEIG (rci90/maneig.f90:157-555) for one block: CSR is the unshifted H
 * as SETHAM writes it; the diagonal is shifted by -eav here.  eval gets
 * nvec eigenvalues of H - EAV, evec ncf x nvec (sign-fixed).  Returns 0, or
 * the DVDSON error code. */
int orc_maneig(int ncf, const int64_t *row_end, const int32_t *col,
               const double *val, double eav, int nvec, double *eval,
               double *evec, int nthreads, int *nloops, int *nmv);

/* MANEIG of rmcdhf90 (rmcdhf90/maneig.f90:78-200), the SCF program's
 * eigensolve: levels iccmin[0..nlev) (1-based) of the block, cold start
 * (INIEST2 with N1000 = 4000) or warm start from the previous eval/evec
 * (in/out), GDVD above 4000 CSFs with LIM = min(NCF, 2*NVECT+40),
 * MAXITR = min(200*NVECT, NCF).  Returns 0 or the DVDSON error code. */
int orc_maneig_mcdhf(int ncf, const int64_t *row_end, const int32_t *col, const double *val,
                     double eav, int nvect, int nlev, const int32_t *iccmin, int warm,
                     double *eval, double *evec, int nthreads, int *nloops, int *nmv);

#ifdef __cplusplus
}
#endif
#endif
