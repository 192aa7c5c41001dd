/* CPU restatement of CRMod/CRTomo's 2.5D forward + sensitivity path.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle the HIP product is checked against, and
 * the "port" CPU baseline.  Nothing in the product links or calls it.
 * Array layouts are those of the reference's Fortran modules (column major,
 * 1-based index VALUES inside integer arrays). */
#ifndef CRT_ORACLE_H
#define CRT_ORACLE_H

typedef struct {
  int sanz, typanz, mb, eanz;
  const int *typ, *nelanz, *selanz; /* typanz */
  const int *nrel;                  /* (elanz+relanz) x maxnodes, col-major */
  const int *rnr;                   /* relanz */
  const int *snr;                   /* sanz */
  const double *sx, *sy;            /* sanz */
  const int *enr;                   /* eanz */
  int lsytop, lsink, nsink;
  int rwdanz;                       /* prescribed boundary values (randb2), 0 = none */
  const int *rwdnr;                 /* rwdanz nodes (1-based) */
  const struct orc_c_ *rwd;         /* rwdanz values (complex; the DC twins take the real part) */
} orc_mesh;

typedef struct orc_c_ { double re, im; } orc_c;

int orc_rwaven(const orc_mesh *m, int nanz, const int *strnr, const int *vnr,
               int *kwnanz, double *kwn, double *kwnwi, double *amin, double *amax);
int orc_precal(const orc_mesh *m, int kwnanz, const double *kwn, double *elbg,
               double *relbg, double *kg, int *lbeta, int *lrandb, double *sytop);
int orc_assemble(const orc_mesh *m, int kwnanz, const double *elbg, const double *relbg,
                 const double *kg, int lrandb, const orc_c *sigma, int l, int k,
                 orc_c *a, orc_c *b, double *fak);
int orc_chol(int sanz, int mb, orc_c *a);
void orc_vre(int sanz, int mb, const orc_c *a, const orc_c *b, orc_c *pot);
int orc_solve_ld(int sanz, int mb, const orc_c *a, const orc_c *b, orc_c *pot);
int orc_forward(const orc_mesh *m, int kwnanz, const double *kwnwi, int swrtr,
                const double *elbg, const double *relbg, const double *kg, int lbeta,
                int lrandb, const orc_c *sigma, orc_c *kpot, orc_c *hpot, int nthreads,
                int sys_begin, int sys_end);
int orc_volt(const orc_mesh *m, int nanz, const int *strnr, const int *vnr,
             const orc_c *hpot, int mode, const double *kfak, orc_c *volt, orc_c *sigmaa);
int orc_sens(const orc_mesh *m, int nanz, const int *strnr, const int *vnr, int kwnanz,
             const double *kwnwi, int swrtr, const double *elbg, const orc_c *kpot,
             int mode, const orc_c *sigma, const orc_c *volt, orc_c *sens, int nthreads);
void orc_update_rhs(int nanz, int manz, const orc_c *sens, const double *wmatd,
                    const int *wdfak, const orc_c *dat, const orc_c *sigmaa, double lam,
                    int ltri, const double *rdiag, const orc_c *rm, double *cgfac,
                    orc_c *bvec);
int orc_cjggra(int nanz, int manz, const orc_c *sens, const double *wmatd, const int *wdfak,
               const double *cgfac, const orc_c *bvec, double lam, int ltri, int nx,
               const double *smatm, int smatm_ld, const int *nachbar, int maxnodes,
               int ncgmax, double eps, orc_c *dpar, double *cgres);
void orc_bsmatmreg(const orc_mesh *m, int nx, int nz, double alfx, double alfz,
                   double *smatm);
void orc_bnachbar(const orc_mesh *m, int *nachbar);
void orc_bsmatmtri(const orc_mesh *m, double alfx, double alfz, const int *nachbar, int ndec,
                   const int *edecoup, const double *edecstr, double *smatm);
void orc_rm_reg(int manz, int nx, const double *smatm, const orc_c *par, orc_c *out);
int orc_assemble_dc(const orc_mesh *m, const double *elbg, const double *relbg, const double *kg,
                    int lrandb, const orc_c *sigma, int l, int k, double *a, double *b, double *fak);
int orc_choldc(int sanz, int mb, double *a);
void orc_vredc(int sanz, int mb, const double *a, const double *b, double *pot);
int orc_forward_dc(const orc_mesh *m, int kwnanz, const double *kwnwi, int swrtr,
                   const double *elbg, const double *relbg, const double *kg, int lbeta,
                   int lrandb, const orc_c *sigma, double *kpotdc, double *hpotdc, int nthreads);
int orc_sens_dc(const orc_mesh *m, int nanz, const int *strnr, const int *vnr, int kwnanz,
                const double *kwnwi, int swrtr, const double *elbg, const double *kpotdc,
                const orc_c *sigma, const orc_c *volt, double *sensdc, int nthreads);
void orc_adjoint_grad(int nanz, int manz, const orc_c *sens, const double *wmatd,
                      const int *wdfak, const orc_c *resid, orc_c *grad);

void orc_chkpo2(orc_c dati, orc_c sigi, int ldum, orc_c *cdat, orc_c *csig, int *wdlok);
int orc_dmisft(int nanz, const orc_c *dat, const orc_c *sigmaa, double *wmatd,
               const double *wmatdr, const double *wmatdp, const orc_c *d0, const orc_c *fm0,
               int flags, double l1min, double *stats, int *wdlok);
void orc_bpar(int elanz, const int *mnr, const orc_c *sigma, orc_c *par);
void orc_bsigma(int elanz, const int *mnr, const orc_c *par, orc_c *sigma);
void orc_update_step(int manz, orc_c *par, const orc_c *dpar, double step, int lphi0, double *out);
double orc_blam0(int nanz, int manz, const orc_c *sens, const double *w, int mode, double alfx,
                 double alfz);

#endif
