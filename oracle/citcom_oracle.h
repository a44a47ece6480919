/*
 * citcom_oracle -- CPU restatement of the CitcomCU timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (citcom_amd/) may link,
 * load or call this library; it is the checker that the tests, the smoke
 * check and the CPU-baseline leg of bench.py compare the HIP path against.
 *
 * Layout conventions (shared with the HIP path so the two compare 1:1):
 *   node  n = k + noz*(i + nox*j)      k = z index, i = x index, j = y index
 *   elem  e = ez + elz*(ex + elx*ey)   (0-based; reference is 1-based)
 *   velocity vectors double[3*nno], dof 3n+d, d = 0:x 1:y 2:z  (= ID doff)
 *   pressure double[nel]; T, Tdot, buoyancy float[nno]
 * Reference arrays that are 1-based (X, NODE, GNX, ...) are stored 0-based.
 */
#ifndef CITCOM_ORACLE_H
#define CITCOM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXL 12

/* node flag bits, global_defs.h:65-89 */
#define ORC_VBX 0x2
#define ORC_VBZ 0x4
#define ORC_VBY 0x8
#define ORC_TBX 0x10
#define ORC_TBZ 0x20
#define ORC_TBY 0x40
#define ORC_FBX 0x80000
#define ORC_FBZ 0x100000
#define ORC_FBY 0x200000

typedef struct orc_params
{
	int levels;	/* levmax = levels-1, levmin = 0 */
	int mgunitx, mgunity, mgunitz;
	int topvbc, botvbc;	/* 0 free slip, 1 prescribed (zero) velocity */
	int toptbc, bottbc;	/* 1 fixed temperature */
	float toptbcval, bottbcval;
	double accuracy;
	int p_iterations;
	int v_steps_low, v_steps_high, down_heavy, up_heavy, mg_cycle;
	int aug_lagr;
	double aug_number;
	int precond;
	/* viscosity law (Viscosity_structures.c:369-443, 491-737) */
	int rheol;	/* 0 or 3 */
	int tdepv;	/* TDEPV */
	int num_mat;
	float visc0[4], viscE[4], viscT[4];
	float z_layer[4];	/* layer tops for num_mat > 1 (layers()) */
	int vmin_on, vmax_on;
	float visc_min, visc_max;
	/* energy */
	float rayleigh;
	float finetunedt;
	int temp_iterations;	/* adv_sub_iterations */
	float gamma;
	double Q0;
	float fixed_timestep;
	/* smoother: 0 = lexicographic node-block GS (reference), 1 = 8-colour
	 * (the coarsest level of multi_grid applied as a matrix, orc_coarse_solve) */
	int smoother;
} orc_params;

typedef struct orc orc;

/* ---- construction ---------------------------------------------------- */
orc *orc_create(const orc_params *p, const float *xc, const float *yc, const float *zc);
void orc_free(orc *o);
/* raw access to an internal array: name as in citcom_oracle.c:orc_get */
void *orc_get(orc *o, const char *name, int lev, long *count);

/* ---- operators (level lev) -------------------------------------------- */
void orc_set_viscosity(orc *o, const float *evi_finest);	/* EVI[levmax], then project + rebuild */
void orc_build_operators(orc *o);	/* project_viscosity + K + BI + EG + BPI on every level */
void orc_matvec(orc *o, int lev, const double *u, double *Au, int strip);
void orc_gauss_seidel(orc *o, int lev, double *d0, const double *F, double *Ad, int cycles, int guess);
/* smoother 1: the coarsest-level smoother of multi_grid as a matrix, d = Mc F */
void orc_coarse_solve(orc *o, const double *F, double *d);
/* smoother 1: the level above the coarsest, out = Ma din + Mb F (din may be NULL) */
void orc_affine_solve(orc *o, const double *din, const double *F, double *out);
void orc_project(orc *o, int lev, const double *fine, double *coarse);
void orc_interp(orc *o, int lev, const double *coarse, double *fine);
void orc_div(orc *o, int lev, const double *U, double *divU);
void orc_grad(orc *o, int lev, const double *P, double *gradP);
void orc_strip(orc *o, int lev, double *v);

/* ---- solvers ------------------------------------------------------------ */
double orc_multi_grid(orc *o, double *d1, double *F, double *Au, double acc, int hl);
int orc_solve_del2_u(orc *o, double *d0, double *F, double acc, int hl, int *count);
/* Uzawa; returns iteration count; U,P,F finest level; V out float[3][nno] */
int orc_solve_stokes(orc *o, double *U, double *P, const double *F, double *stats);

/* ---- energy and forces -------------------------------------------------- */
void orc_thermal_buoyancy(orc *o, const float *T, float *buoy);
void orc_assemble_forces(orc *o, const float *buoy, double *F);
void orc_visc_from_T(orc *o, const float *T, float *evi);
/* one PG_timestep; T, Tdot in/out; V in (float[3][nno]); returns dt */
float orc_pg_timestep(orc *o, float *T, float *Tdot, const float *V, float *stats);
double orc_vrms(orc *o, const float *V);	/* volume-RMS velocity (no reference routine) */
void orc_heat_flux(orc *o, const float *T, const float *V, float *nu);

#ifdef __cplusplus
}
#endif
#endif
