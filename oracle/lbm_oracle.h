/*
 * lbm_oracle.h -- CPU oracle of the reference LBM time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ORC_LEX_WAVE: the results of ORC_LEX_GS computed in parallel (D3Q19, without the
 * interpolated bounce-back, whose diagonal neighbour reads it cannot order) */
enum { ORC_JACOBI = 0, ORC_LEX_GS = 1, ORC_LEX_WAVE = 2 };
enum { ORC_SGS_WALE = 0, ORC_SGS_SMAGORINSKY = 1, ORC_SGS_NONE = 2 };
/* D3Q19: RR (the reference's) or BGK; D3Q27: RR (selectable) or CUMULANT (the reference's) */
enum { ORC_COLL_RR = 0, ORC_COLL_BGK = 1, ORC_COLL_CUMULANT = 2 };

typedef struct {
    int nx, ny, nz;
    int order;     /* ORC_JACOBI | ORC_LEX_GS */
    int sgs;       /* ORC_SGS_* */
    int collision; /* ORC_COLL_* */
    int damping;   /* Smagorinsky wall damping: 0 off, 1 Fluent, 2 van Driest */
    float dpdx, LES, omegaB, spzWidth, Cs;
    float clim;    /* D3Q27 cumulant limiter (input.hpp default 0.01) */
    int ibb;       /* 1: interpolated bounce-back with the precomputed q-distances below */
    const void *qf; /* q-distances, Q*N reals of the step's precision (float or double), q-major */
} orc_params;

extern const int ORC_CX[19], ORC_CY[19], ORC_CZ[19];

void orc_fatal(const char *msg);
long orc_index(int i, int j, int k, int nx, int ny);
int orc_reflect(int q);
int orc_reflect_or_mirror(int q, int b1, int b2, int b3);
long orc_upwind_B(int q, int i, int j, int k, int nx, int ny, int nz, int b1, int b2, int b3);
long orc_upwind_internal(int q, int i, int j, int k, int nx, int ny);
long orc_upwind_2D(int q, int i, int j, int k, int nx, int ny, int nz);
int orc_corner(int b3, int i, int j, int k, int nx, int ny, int nz);

int orc19_step_f32(const orc_params *P, float *f, float *fTmp, const int *boundary, const int *b1, const int *b2,
                   const int *b3, const float *sdf, const unsigned char *solid, const unsigned char *neiSolid,
                   const float *u0, const float *v0, const float *w0, float *Fwx, float *Fwy, float *Fwz,
                   float *tauSGS, float *rho, float *u, float *v, float *w, const float *omega, float *Gx,
                   float *Gy, float *Gz);
int orc19_step_f64(const orc_params *P, double *f, double *fTmp, const int *boundary, const int *b1, const int *b2,
                   const int *b3, const double *sdf, const unsigned char *solid, const unsigned char *neiSolid,
                   const double *u0, const double *v0, const double *w0, double *Fwx, double *Fwy, double *Fwz,
                   double *tauSGS, double *rho, double *u, double *v, double *w, const double *omega, double *Gx,
                   double *Gy, double *Gz);
void orc19_collide_f32(int kind, const float *ft, float rho, float u, float v, float w, float tau, float tauSGS,
                       float omegaB, float dpdx, float *out);
void orc19_collide_f64(int kind, const double *ft, double rho, double u, double v, double w, double tau,
                       double tauSGS, double omegaB, double dpdx, double *out);
float orc19_smagorinsky_f32(const float *ft, float rho, float u, float v, float w, float omega, float LES, float sdf,
                            float Cs, int damping);
double orc19_smagorinsky_f64(const double *ft, double rho, double u, double v, double w, double omega, double LES,
                             double sdf, double Cs, int damping);
/* van Driest damping factor of tauSGS (damping 2) at wall distance y and speed U; y+ out */
float orc_van_driest_f32(float tau, float U, float y, float Cs, float *yplus);
double orc_van_driest_f64(double tau, double U, double y, double Cs, double *yplus);

/* D3Q27 */
extern const int ORC27_CX[27], ORC27_CY[27], ORC27_CZ[27];
int orc27_reflect(int q);
int orc27_mirror(int q, int axis);
long orc27_upwind_B(int q, int i, int j, int k, int nx, int ny, int nz, int b1, int b2, int b3);
int orc27_step_f32(const orc_params *P, float *f, float *fTmp, const int *boundary, const int *b1, const int *b2,
                   const int *b3, const float *sdf, const unsigned char *solid, const unsigned char *neiSolid,
                   const float *u0, const float *v0, const float *w0, float *Fwx, float *Fwy, float *Fwz,
                   float *tauSGS, float *rho, float *u, float *v, float *w, const float *omega, float *Gx,
                   float *Gy, float *Gz);
int orc27_step_f64(const orc_params *P, double *f, double *fTmp, const int *boundary, const int *b1, const int *b2,
                   const int *b3, const double *sdf, const unsigned char *solid, const unsigned char *neiSolid,
                   const double *u0, const double *v0, const double *w0, double *Fwx, double *Fwy, double *Fwz,
                   double *tauSGS, double *rho, double *u, double *v, double *w, const double *omega, double *Gx,
                   double *Gy, double *Gz);
void orc27_collide_f32(const float *ft, float rho, float u, float v, float w, float tau, float tauSGS, float omegaB,
                       float clim, float *out);
void orc27_collide_f64(const double *ft, double rho, double u, double v, double w, double tau, double tauSGS,
                       double omegaB, double clim, double *out);
void orc27_collide_rr_f32(const float *ft, float rho, float u, float v, float w, float tau, float tauSGS,
                          float omegaB, float dpdx, float *out);
void orc27_collide_rr_f64(const double *ft, double rho, double u, double v, double w, double tau, double tauSGS,
                          double omegaB, double dpdx, double *out);

/* immersed boundary (moving walls); m[10] = trans[3], rot centre[3], rot axis[3], omega */
long orc_ibm_icbox(long ic, int b, int nx, int ny, int nz);
void orc_ibm_urot_f32(const float *m, float x, float y, float z, float *ur, float *vr, float *wr);
void orc_ibm_xrot_f32(const float *m, float x, float y, float z, float *X, float *Y, float *Z);
void orc_ibm_gwall_f32(const float *m, int n, const float *xyz, int nx, int ny, int nz, const float *u,
                       const float *v, const float *w, float *Gx, float *Gy, float *Gz);
void orc_ibm_spread_move_f32(const float *m, int n, float *xyz, int nx, int ny, int nz, const float *Gmx,
                             const float *Gmy, const float *Gmz, float *Gx, float *Gy, float *Gz);
void orc_ibm_force_f32(int Q, long N, float *f, const unsigned char *solid, const float *rho, const float *Gx,
                       const float *Gy, const float *Gz);
void orc_ibm_urot_f64(const double *m, double x, double y, double z, double *ur, double *vr, double *wr);
void orc_ibm_xrot_f64(const double *m, double x, double y, double z, double *X, double *Y, double *Z);
void orc_ibm_gwall_f64(const double *m, int n, const double *xyz, int nx, int ny, int nz, const double *u,
                       const double *v, const double *w, double *Gx, double *Gy, double *Gz);
void orc_ibm_spread_move_f64(const double *m, int n, double *xyz, int nx, int ny, int nz, const double *Gmx,
                             const double *Gmy, const double *Gmz, double *Gx, double *Gy, double *Gz);
void orc_ibm_force_f64(int Q, long N, double *f, const unsigned char *solid, const double *rho, const double *Gx,
                       const double *Gy, const double *Gz);

/* geometry preprocessing (geom_oracle.c): walls.hpp:18-184 setSolid, walls.hpp:218-297 setSDF */
int orc_geom_is_cross(const float *v0, const float *v1, const float *v2, int i, int j, int k);
int orc_geom_solid(int n, const float *v0, const float *v1, const float *v2, int nx, int ny, int nz, int invert,
                   unsigned char *solid);
int orc_geom_sdf(int n, const float *c, const float *nrm, float sdfIni, float dr, float p, int nx, int ny, int nz,
                 const unsigned char *neiSolid, float *sdf);
int orc_geom_qlinks(int Q, int n, const float *v0, const float *v1, const float *v2, int nx, int ny, int nz,
                    const unsigned char *solid, const unsigned char *neiSolid, float *qlink, int *misses);

/* output step (output_oracle.c): cal_rhoUVW, D3Q19_opencl/D3Q19.hpp:647-666 */
void orc_cal_rhouvw(int Q, long N, const float *f, float dpdx, float *rho, float *u, float *v, float *w);

#ifdef __cplusplus
}
#endif
#endif
