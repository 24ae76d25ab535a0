/*
 * pda_oracle.h -- CPU restatement of pressio-demoapps' cell-centered
 * finite-volume residual (velocity) and Jacobian assembly.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker that the HIP
 * product is compared against.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never links it.
 *
 * The algorithm follows the reference files cited per function in
 * pda_oracle.c (paths relative to the pressio-demoapps source tree).
 */
#ifndef PDA_ORACLE_H
#define PDA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* problem families (same numbering as the product C-ABI) */
enum {
  ORC_EULER1D = 1, ORC_EULER2D = 2, ORC_EULER3D = 3,
  ORC_SWE2D = 4, ORC_BURGERS2D = 5, ORC_ADR2D = 6
};

/* reconstruction: include/pressiodemoapps/schemes_info.hpp:61-63 */
enum { ORC_FIRSTORDER = 0, ORC_WENO3 = 1, ORC_WENO5 = 2 };

typedef struct {
  int dim;            /* 1,2,3 */
  int stencil;        /* mesh stencil size 3,5,7 */
  int64_t nrows;      /* sample mesh size (graph rows) */
  int64_t ncells;     /* stencil mesh size (state cells) */
  const int32_t *graph; /* nrows x (1+(stencil-1)*dim), row-major */
  double dx[3];       /* cell widths as read from info.dat */
  const double *x, *y, *z; /* cell-centre coordinates by stencil-mesh id */
} orc_mesh;

typedef struct {
  int family;
  int prob;           /* reference enum value within the family */
  int scheme;         /* ORC_FIRSTORDER / ORC_WENO3 / ORC_WENO5 */
  /* physical parameters:
     Euler*: p[0]=gamma;  SWE: p[0]=gravity p[1]=coriolis;
     Burgers: p[0]=diffusion;  ADR: p[0]=ux p[1]=uy p[2]=diffusion p[3]=sigma */
  double p[8];
} orc_problem;

/* returns number of dofs per cell for the family, or -1 */
int orc_ndpc(int family);

/* velocity-only evaluation (reference rightHandSide).  V has nrows*ndpc. */
int orc_velocity(const orc_mesh *mesh, const orc_problem *prob,
                 const double *U, double t, double *V);

/* Jacobian sparsity: scalar CSR with sorted, merged columns, same as Eigen
   setFromTriplets + makeCompressed.  Call with rowptr==NULL to get nnz. */
int orc_jacobian_pattern(const orc_mesh *mesh, const orc_problem *prob,
                         int64_t *nnz, int64_t *rowptr, int32_t *cols);

/* velocity-and-Jacobian evaluation (reference rightHandSideAndJacobian).
   vals are in the order given by orc_jacobian_pattern. */
int orc_velocity_and_jacobian(const orc_mesh *mesh, const orc_problem *prob,
                              const double *U, double t, double *V,
                              const int64_t *rowptr, const int32_t *cols,
                              double *vals);

/* scalar kernels exposed for known-answer tests */
void orc_weno5(double *uNeg, double *uPos, const double q[6]);
void orc_weno5_grad(double *uNeg, double *uPos, double dNeg[6], double dPos[6],
                    const double q[6]);
void orc_weno3(double *uNeg, double *uPos, const double q[4]);
void orc_weno3_grad(double *uNeg, double *uPos, double dNeg[4], double dPos[4],
                    const double q[4]);
/* flux and flux Jacobians at one face; n = unit normal (dim entries) */
int orc_face_flux(const orc_problem *prob, int dim, const double *qL,
                  const double *qR, const double *n, double *F);
int orc_face_flux_jac(const orc_problem *prob, int dim, const double *qL,
                      const double *qR, const double *n, double *JL, double *JR);

#ifdef __cplusplus
}
#endif
#endif
