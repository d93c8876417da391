/*
 * CPU oracle for the Vanadis element-by-element DR-HEX8 operator and BiCG.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the reference the GPU results are compared with
 * (or the timed CPU baseline).  The product (libvanadis_gpu.so) never links it.
 *
 * Every routine restates one reference routine in plain C, in the reference's
 * loop and operation order (built with -ffp-contract=off, like gfortran on
 * baseline x86-64 which emits no FMA).  Citations are to the reference tree:
 *   V1 = src_v2026_2_1/VANADIS.FOR            (release, CUDA-linked)
 *   V2 = src_updates/gauss_point_hex8DR.for    (standalone update)
 *   V3 = src_updates/vanadis_full_DR_OpenMP_timing.for (CPU OpenMP timing build)
 *
 * Conventions: element matrices are row-major, A[8*i + j] = A(i+1, j+1) of the
 * Fortran 8x8 arrays.  ISS is column-major like the Fortran ISS(E,8):
 * iss[j*E + e] is the 1-based node of local node j+1 of element e+1.  KOD is
 * kod[f*E + e].  Per-element fields are SoA: vel[e], vel[E+e], vel[2E+e].
 *
 * Parity pins (tests/): reference run transcript values (DDF, masses,
 * iterations), the flang-built reference itself (oracle/_ref, when built),
 * exact Kronecker closed forms for box elements, scipy CSR/spsolve/cg.
 */
#ifndef VANADIS_ORACLE_H
#define VANADIS_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_V1 = 1, ORC_V2 = 2, ORC_V3 = 3 };
/* element-system schemes */
enum { ORC_THETA = 0, ORC_STEADY = 1, ORC_EULER = 2 };
/* Dirichlet face maps: V1 (faces 5/6 = y-max/y-min), V3 (swapped) */
enum { ORC_FACEMAP_V1 = 1, ORC_FACEMAP_V3 = 3 };

void orc_siatka(const double *dx, const double *dy, const double *dz, int nax,
                int nay, int naz, double *xx, double *yy, double *zz);
void orc_w_el(int nax, int nay, int naz, int *iss);
void orc_kod_el_dr(int nax, int nay, int naz, int *kod);
void orc_build_bc_nodes(int facemap, const int *iss, const int *kod, int nax,
                        int nay, int naz, int *mask);

void orc_shape_hex8(double xi, double eta, double zeta, double n[8],
                    double dn[24]);
void orc_quad4_3d(const double x[4], const double y[4], const double z[4],
                  double r, double s, double *detjs);
void orc_gauss_point(int variant, double xi, double eta, double zeta,
                     const double xn[8], const double yn[8], const double zn[8],
                     const double lam[3], const double v[3], double *detj,
                     double n[8], double w[8], double dnx[8], double dny[8],
                     double dnz[8], double *tau_out);
void orc_el3d(int variant, const double xe[8], const double ye[8],
              const double ze[8], const double lam[3], double P,
              const double v[3], double q_in, int robin, double alpha,
              double qs, double H[64], double C[64], double F[8]);

int orc_assemble(int variant, int scheme, double dt, int nax, int nay, int naz,
                 const double *xx, const double *yy, const double *zz,
                 const int *iss, const int *kod, const double *lam,
                 double P, const double *vel, const double *q_in, double alpha,
                 double qs, const double *t_old, double *bb, double *fmv3,
                 double *diag);

void orc_matvec_0(double *xin, double *xout, const int *iss, const double *bb,
                  int nax, int nay, int naz, const int *mask);
/* matvec_0 over 8 parity colours of elements, OpenMP within a colour */
int orc_matvec_0_colored(double *xin, double *xout, const int *iss, const double *bb,
                         int nax, int nay, int naz, const int *mask, int threads);
void orc_mttvec_0(double *xin, double *xout, const int *iss, const double *bb,
                  int nax, int nay, int naz, const int *mask);
double orc_dnrm2(int n, const double *x);
int orc_dbcg_simple(int n, const double *b, double *x, double tol, int itmax,
                    int *iter, double *err, const double *bb, const int *iss,
                    const double *diag, int nax, int nay, int naz,
                    const int *mask);

/* OpenMP CPU path (ompbase.c): mttvec_0 over the 8 parity colours; the
 * reference BiCG with colour-parallel products and fixed-order OpenMP
 * reductions (bitwise reproducible for a fixed thread count); the assembly
 * with parallel element systems (bitwise orc_assemble). */
int orc_mttvec_0_colored(double *xin, double *xout, const int *iss, const double *bb,
                         int nax, int nay, int naz, const int *mask, int threads);
int orc_dbcg_simple_omp(int n, const double *b, double *x, double tol, int itmax,
                        int *iter, double *err, const double *bb, const int *iss,
                        const double *diag, int nax, int nay, int naz, const int *mask,
                        int threads);
int orc_assemble_omp(int variant, int scheme, double dt, int nax, int nay, int naz,
                     const double *xx, const double *yy, const double *zz,
                     const int *iss, const int *kod, const double *lam, double P,
                     const double *vel, const double *q_in, double alpha, double qs,
                     const double *t_old, double *bb, double *fmv3, double *diag,
                     int threads);

double orc_integrate_mass_hex8(int nelem, const int *iss, const double *xx,
                               const double *yy, const double *zz,
                               const double *c);
double orc_diagonal_dominance_factor(int nelem, int nnode, const int *iss,
                                     const double *bb);
int orc_is_diagonally_dominant(int nelem, int nnode, const int *iss,
                               const double *bb);

void orc_generate_hex8_colors(int nax, int nay, int naz, int *color);
void orc_build_color_element_lists(int ile, const int *color, int *elems,
                                   int offsets[9], int counts[8]);

#ifdef __cplusplus
}
#endif
#endif
