/*
 * citcom_oracle.c -- CPU restatement of the CitcomCU timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see citcom_oracle.h).  Each function follows the
 * reference routine cited beside it, loop for loop, keeping the reference's
 * float/double rounding points (float storage of higher_precision data,
 * double accumulators) so that results can be compared bit for bit where the
 * reference is deterministic.  Built with -ffp-contract=off for that reason.
 *
 * Paths below are relative to the reference's src/ directory.
 */
#include "citcom_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define BETA 0.57735026918962576451
#define MAXL ORC_MAXL
#define fmaxf_(a, b) (((a) > (b)) ? (a) : (b))
#define fminf_(a, b) (((a) < (b)) ? (a) : (b))

/* per-element record sizes (global_defs.h:240-269, 290-293, 313-319) */
#define GNX_SZ 216	/* Shape_function_dx: vpt[3*8*8] + ppt[3*8] */
#define GDA_SZ 11	/* Shape_function_dA: vpt[9] + ppt[2] */
#define ECO_SZ 13	/* COORD: centre[4] size[4] recip_size[4] area */
#define TWW_SZ 9	/* FNODE: node[9] */
#define EG_SZ 24	/* EG: g[24][1] */
#define KROW 243	/* 3 rows x 27 neighbour nodes x 3 columns */

#define GNV(n, v) (8 * ((n) - 1) + ((v) - 1))	/* GNVINDEX, element_definitions.h:59 */
#define GNVX(d, n, v) (64 * (d) + 8 * ((n) - 1) + ((v) - 1))	/* GNVXINDEX :62 */
#define GNPX(d, n) (8 * (d) + ((n) - 1))	/* GNPXINDEX with p = 1 */

/* element_definitions.h:171-179 (g_point), :214-222 (offset), :231-245 (bb) */
static const float gpx[9][3] = { {0, 0, 0},
	{-(float)BETA, -(float)BETA, -(float)BETA}, {(float)BETA, -(float)BETA, -(float)BETA},
	{(float)BETA, (float)BETA, -(float)BETA}, {-(float)BETA, (float)BETA, -(float)BETA},
	{-(float)BETA, -(float)BETA, (float)BETA}, {(float)BETA, -(float)BETA, (float)BETA},
	{(float)BETA, (float)BETA, (float)BETA}, {-(float)BETA, (float)BETA, (float)BETA} };
/* local node a -> (dz, dx, dy) offsets relative to local node 1 */
static const int offz[9] = { 0, 0, 0, 0, 0, 1, 1, 1, 1 };
static const int offx[9] = { 0, 0, 1, 1, 0, 0, 1, 1, 0 };
static const int offy[9] = { 0, 0, 0, 1, 1, 0, 0, 1, 1 };
static const int bb[3][9] = { {0, 1, 2, 2, 1, 1, 2, 2, 1}, {0, 1, 1, 2, 2, 1, 1, 2, 2}, {0, 1, 1, 1, 1, 2, 2, 2, 2} };
static const int epsilon_[4][4] = { {0, 0, 0, 0}, {0, 1, -1, 1}, {0, -1, 1, -1}, {0, 1, -1, 1} };
/* element_definitions.h:187-191 (s_point x,y for the 4 surface points) */
static const float spx[5][2] = { {0, 0}, {-(float)BETA, -(float)BETA}, {(float)BETA, -(float)BETA},
	{(float)BETA, (float)BETA}, {-(float)BETA, (float)BETA} };

struct orc
{
	orc_params p;
	int levmin, levmax;
	int nox[MAXL], noy[MAXL], noz[MAXL], elx[MAXL], ely[MAXL], elz[MAXL];
	int nno[MAXL], nel[MAXL], neq[MAXL];
	float *X[MAXL][3];
	unsigned *NODE[MAXL];
	float *GNX[MAXL], *GDA[MAXL], *ECO[MAXL], *TWW[MAXL], *MASS[MAXL];
	float *EVI[MAXL], *EG[MAXL], *K[MAXL];
	double *BI[MAXL], *BPI[MAXL];
	int *mat;	/* finest level, 1-based material id per element */
	float *TB[3];	/* finest level */
	float *expansivity, *diffusivity;	/* [noz] */
	/* master element (Shape_functions.c:53-170) */
	double Nv[64], Np[8];
	float Nxv[192], Nxp[24];
	double Mv[16], Mxv[32];
	/* energy state */
	float diff_timestep;
	int dt_been_here;
	float elapsed_time;
	/* work */
	double *mg_vel[MAXL], *mg_res[MAXL], *mg_rhs[MAXL], *mg_fl[MAXL], *mg_del[MAXL], *mg_AU[MAXL];
	double *Mc;	/* coarsest-level smoother as a matrix (smoother 1), column-major */
	double *Ma, *Mb;	/* level levmin+1: its smoother call d = Ma din + Mb F (smoother 1), column-major */
};

static double lpoly(int p, double y)
{
	return p == 1 ? 0.5 * (1 - y) : (p == 2 ? 0.5 * (1 + y) : 0.0);
}
static double lpolydash(int p, double y)
{
	(void)y;
	return p == 1 ? -0.5 : (p == 2 ? 0.5 : 0.0);
}

/* construct_shape_functions, Shape_functions.c:53-170 (3-D, cartesian) */
static void shape_functions(orc *o)
{
	int i, j, d, dd, k;
	for(i = 1; i <= 8; i++)
	{
		for(j = 1; j <= 8; j++)
		{
			o->Nv[GNV(i, j)] = 1.0;
			for(d = 1; d <= 3; d++)
				o->Nv[GNV(i, j)] *= lpoly(bb[d - 1][i], gpx[j][d - 1]);
			for(dd = 1; dd <= 3; dd++)
			{
				float v = (float)lpolydash(bb[dd - 1][i], gpx[j][dd - 1]);
				for(d = 1; d <= 3; d++)
					if(d != dd)
						v = (float)((double)v * lpoly(bb[d - 1][i], gpx[j][d - 1]));
				o->Nxv[GNVX(dd - 1, i, j)] = v;
			}
		}
		/* the single pressure point at the element centre */
		o->Np[i - 1] = 1.0;
		for(d = 1; d <= 3; d++)
			o->Np[i - 1] *= lpoly(bb[d - 1][i], 0.0);
		for(dd = 1; dd <= 3; dd++)
		{
			float v = (float)lpolydash(bb[dd - 1][i], 0.0);
			for(d = 1; d <= 3; d++)
				if(d != dd)
					v = (float)((double)v * lpoly(bb[d - 1][i], 0.0));
			o->Nxp[GNPX(dd - 1, i)] = v;
		}
	}
	/* 2-D surface functions M, Mx on the 4 surface points (GMVINDEX) */
	for(j = 1; j <= 4; j++)
		for(k = 1; k <= 4; k++)
		{
			o->Mv[4 * (j - 1) + k - 1] = lpoly(bb[0][j], spx[k][0]) * lpoly(bb[1][j], spx[k][1]);
			o->Mxv[0 * 16 + 4 * (j - 1) + k - 1] = lpolydash(bb[0][j], 0) * lpoly(bb[1][j], spx[k][1]);
			o->Mxv[1 * 16 + 4 * (j - 1) + k - 1] = lpolydash(bb[1][j], 0) * lpoly(bb[0][j], spx[k][0]);
		}
}

static inline int nidx(const orc *o, int lev, int k, int i, int j)
{
	return k + o->noz[lev] * (i + o->nox[lev] * j);
}

/* node of local index a (1..8) of element (ez,ex,ey) at level lev */
static inline int enode(const orc *o, int lev, int ez, int ex, int ey, int a)
{
	return nidx(o, lev, ez + offz[a], ex + offx[a], ey + offy[a]);
}

static inline void edecomp(const orc *o, int lev, int e, int *ez, int *ex, int *ey)
{
	*ez = e % o->elz[lev];
	*ex = (e / o->elz[lev]) % o->elx[lev];
	*ey = e / (o->elz[lev] * o->elx[lev]);
}

static double det3(double A[4][4])
{
	return (A[1][1] * (A[2][2] * A[3][3] - A[2][3] * A[3][2]) - A[1][2] * (A[2][1] * A[3][3] - A[2][3] * A[3][1]) + A[1][3] * (A[2][1] * A[3][2] - A[2][2] * A[3][1]));
}

/* cofactor, General_matrix_functions.c:1373-1404 (n = 3) */
static double cofactor3(double A[4][4], int i, int j)
{
	double B[4][4];
	int k, l, p = 1, q = 1;
	for(k = 1; k <= 3; k++)
	{
		if(k == i)
			continue;
		for(l = 1; l <= 3; l++)
		{
			if(l == j)
				continue;
			B[p][q] = A[k][l];
			q++;
		}
		q = 1;
		p++;
	}
	return epsilon_[i][j] * (B[1][1] * B[2][2] - B[1][2] * B[2][1]);
}

/* get_global_shape_fn, Size_does_matter.c:60-179, cartesian, both the
 * velocity points (pressure = 0) and the pressure point (pressure = 2) */
static void global_shape_fn(orc *o, int lev, int el)
{
	int ez, ex, ey, i, j, k, d, e;
	int nodes[9];
	double dxda[4][4], cof[4][4], jac;
	float *gnx = o->GNX[lev] + (size_t)el * GNX_SZ, *gda = o->GDA[lev] + (size_t)el * GDA_SZ;
	edecomp(o, lev, el, &ez, &ex, &ey);
	for(i = 1; i <= 8; i++)
		nodes[i] = enode(o, lev, ez, ex, ey, i);
	for(k = 1; k <= 9; k++)
	{			/* k = 9 stands for the pressure point */
		const int kk = (k == 9) ? 1 : k;
		for(d = 1; d <= 3; d++)
			for(e = 1; e <= 3; e++)
				dxda[d][e] = 0.0;
		for(i = 1; i <= 8; i++)
			for(d = 1; d <= 3; d++)
				for(e = 1; e <= 3; e++)
				{
					float nx = (k == 9) ? o->Nxp[GNPX(d - 1, i)] : o->Nxv[GNVX(d - 1, i, kk)];
					float pr = o->X[lev][e - 1][nodes[i]] * nx;
					dxda[d][e] += pr;
				}
		jac = det3(dxda);
		if(k == 9)
			gda[9 + 1] = (float)jac;
		else
			gda[k] = (float)jac;
		for(d = 1; d <= 3; d++)
			for(e = 1; e <= 3; e++)
				cof[d][e] = cofactor3(dxda, d, e);
		for(j = 1; j <= 8; j++)
			for(d = 1; d <= 3; d++)
			{
				float g = 0.0f;
				for(e = 1; e <= 3; e++)
				{
					float nx = (k == 9) ? o->Nxp[GNPX(e - 1, j)] : o->Nxv[GNVX(e - 1, j, kk)];
					g = (float)((double)g + nx * cof[e][d]);
				}
				g = (float)((double)g / jac);
				if(k == 9)
					gnx[192 + GNPX(d - 1, j)] = g;
				else
					gnx[GNVX(d - 1, j, kk)] = g;
			}
	}
}

/* mass_matrix, Size_does_matter.c:618-757, cartesian */
static void mass_matrix(orc *o, int lev)
{
	int e, a, n, nint;
	int nodes[9];
	for(n = 0; n < o->nno[lev]; n++)
		o->MASS[lev][n] = 0.0f;
	for(e = 0; e < o->nel[lev]; e++)
	{
		int ez, ex, ey, i;
		double area = 0.0, centre[4] = { 0, 0, 0, 0 }, temp[9];
		float dx1, dx2, dx3;
		float *eco = o->ECO[lev] + (size_t)e * ECO_SZ;
		const float *gda = o->GDA[lev] + (size_t)e * GDA_SZ;
		float *const *XX = o->X[lev];
		edecomp(o, lev, e, &ez, &ex, &ey);
		for(a = 1; a <= 8; a++)
			nodes[a] = enode(o, lev, ez, ex, ey, a);
		for(i = 1; i <= 3; i++)
		{
			for(a = 1; a <= 8; a++)
				centre[i] += XX[i - 1][nodes[a]];
			eco[i] = (float)(centre[i] / 8);
		}
		dx2 = (float)(0.25 * (XX[1][nodes[3]] + XX[1][nodes[4]] + XX[1][nodes[7]] + XX[1][nodes[8]] - XX[1][nodes[1]] - XX[1][nodes[2]] - XX[1][nodes[5]] - XX[1][nodes[6]]));
		eco[4 + 2] = (float)sqrt(dx2 * dx2);
		dx1 = (float)(0.25 * (XX[0][nodes[2]] + XX[0][nodes[3]] + XX[0][nodes[6]] + XX[0][nodes[7]] - XX[0][nodes[1]] - XX[0][nodes[4]] - XX[0][nodes[5]] - XX[0][nodes[8]]));
		eco[4 + 1] = (float)sqrt(dx1 * dx1);
		dx3 = (float)(0.25 * (XX[2][nodes[5]] + XX[2][nodes[6]] + XX[2][nodes[7]] + XX[2][nodes[8]] - XX[2][nodes[1]] - XX[2][nodes[2]] - XX[2][nodes[3]] - XX[2][nodes[4]]));
		eco[4 + 3] = (float)sqrt(dx3 * dx3);
		for(nint = 1; nint <= 8; nint++)
		{
			float w = 1.0f * gda[nint];
			area += w;
		}
		eco[12] = (float)area;
		for(a = 1; a <= 8; a++)
		{
			temp[a] = 0.0;
			for(nint = 1; nint <= 8; nint++)
			{
				float w = gda[nint] * 1.0f;
				temp[a] += w * o->Nv[GNV(a, nint)];
			}
		}
		for(a = 1; a <= 8; a++)
			o->MASS[lev][nodes[a]] = (float)((double)o->MASS[lev][nodes[a]] + temp[a]);
		for(a = 1; a <= 8; a++)
			o->TWW[lev][(size_t)e * TWW_SZ + a] = (float)temp[a];
	}
	for(n = 0; n < o->nno[lev]; n++)
		o->MASS[lev][n] = (float)(1.0 / o->MASS[lev][n]);
}

/* velocity_boundary_conditions + velocity_refl_vert_bc + horizontal_bc,
 * Boundary_conditions.c:44-156, 249-372, 602-643; temperature flags
 * temperature_boundary_conditions + temperature_refl_vert_bc :188-246, :516-565 */
static void boundary_flags(orc *o)
{
	int lev, i, j, k;
	for(lev = o->levmin; lev <= o->levmax; lev++)
	{
		const int nox = o->nox[lev], noy = o->noy[lev], noz = o->noz[lev];
		unsigned *NODE = o->NODE[lev];
		for(j = 0; j < noy; j++)
			for(i = 0; i < nox; i++)
				for(k = 0; k < noz; k++)
				{
					unsigned f = 0;
					const int zb = (k == 0 || k == noz - 1);
					/* top / bottom free slip: VBZ on, VBX/VBY off (horizontal_bc) */
					if(zb)
						f |= ORC_VBZ;
					/* x walls (refl_vert_bc): VBX on; interior z: VBY, VBZ off */
					if(i == 0 || i == nox - 1)
					{
						f |= ORC_VBX;
						if(!zb)
							f &= ~(unsigned)(ORC_VBY | ORC_VBZ);
					}
					/* y walls: VBY on; interior z: VBZ off; interior x and z: VBX off */
					if(j == 0 || j == noy - 1)
					{
						f |= ORC_VBY;
						if(!zb)
							f &= ~(unsigned)ORC_VBZ;
						if(i != 0 && i != nox - 1 && !zb)
							f &= ~(unsigned)ORC_VBX;
					}
					/* prescribed velocity top / bottom (second loop) */
					if((k == 0 && o->p.botvbc == 1) || (k == noz - 1 && o->p.topvbc == 1))
						f |= ORC_VBX | ORC_VBY | ORC_VBZ;
					NODE[nidx(o, lev, k, i, j)] = f;
				}
	}
	/* temperature flags on the finest level only */
	lev = o->levmax;
	for(j = 0; j < o->noy[lev]; j++)
		for(i = 0; i < o->nox[lev]; i++)
			for(k = 0; k < o->noz[lev]; k++)
			{
				const int n = nidx(o, lev, k, i, j);
				if(k == 0 && o->p.bottbc >= 1)
				{
					o->NODE[lev][n] |= ORC_TBZ;
					o->TB[2][n] = o->p.bottbcval;
				}
				if(k == o->noz[lev] - 1 && o->p.toptbc >= 1)
				{
					o->NODE[lev][n] |= ORC_TBZ;
					o->TB[2][n] = o->p.toptbcval;
				}
				if(i == 0 || i == o->nox[lev] - 1)
					o->NODE[lev][n] |= ORC_FBX;
				if(j == 0 || j == o->noy[lev] - 1)
					o->NODE[lev][n] |= ORC_FBY;
			}
}

orc *orc_create(const orc_params *p, const float *xc, const float *yc, const float *zc)
{
	orc *o = calloc(1, sizeof *o);
	int lev, i, j, k, e, d;
	o->p = *p;
	o->levmin = 0;
	o->levmax = p->levels - 1;
	shape_functions(o);
	for(lev = o->levmin; lev <= o->levmax; lev++)
	{
		const int f = 1 << lev;
		o->nox[lev] = p->mgunitx * f + 1;
		o->noy[lev] = p->mgunity * f + 1;
		o->noz[lev] = p->mgunitz * f + 1;
		o->elx[lev] = o->nox[lev] - 1;
		o->ely[lev] = o->noy[lev] - 1;
		o->elz[lev] = o->noz[lev] - 1;
		o->nno[lev] = o->nox[lev] * o->noy[lev] * o->noz[lev];
		o->nel[lev] = o->elx[lev] * o->ely[lev] * o->elz[lev];
		o->neq[lev] = 3 * o->nno[lev];
		for(d = 0; d < 3; d++)
			o->X[lev][d] = malloc(sizeof(float) * o->nno[lev]);
		o->NODE[lev] = calloc(o->nno[lev], sizeof(unsigned));
		o->GNX[lev] = calloc((size_t)o->nel[lev] * GNX_SZ, sizeof(float));
		o->GDA[lev] = calloc((size_t)o->nel[lev] * GDA_SZ, sizeof(float));
		o->ECO[lev] = calloc((size_t)o->nel[lev] * ECO_SZ, sizeof(float));
		o->TWW[lev] = calloc((size_t)o->nel[lev] * TWW_SZ, sizeof(float));
		o->MASS[lev] = calloc(o->nno[lev], sizeof(float));
		o->EVI[lev] = calloc((size_t)o->nel[lev] * 8, sizeof(float));
		o->EG[lev] = calloc((size_t)o->nel[lev] * EG_SZ, sizeof(float));
		o->K[lev] = calloc((size_t)o->nno[lev] * KROW, sizeof(float));
		o->BI[lev] = calloc(o->neq[lev], sizeof(double));
		o->BPI[lev] = calloc(o->nel[lev], sizeof(double));
		o->mg_vel[lev] = calloc(o->neq[lev], sizeof(double));
		o->mg_res[lev] = calloc(o->neq[lev], sizeof(double));
		o->mg_rhs[lev] = calloc(o->neq[lev], sizeof(double));
		o->mg_fl[lev] = calloc(o->neq[lev], sizeof(double));
		o->mg_del[lev] = calloc(o->neq[lev], sizeof(double));
		o->mg_AU[lev] = calloc(o->neq[lev], sizeof(double));
	}
	/* coordinates: finest from the per-axis arrays (node_locations,
	 * Nodal_mesh.c:170-180), coarser by injection (:244-247) */
	for(lev = o->levmin; lev <= o->levmax; lev++)
	{
		const int s = 1 << (o->levmax - lev);
		for(j = 0; j < o->noy[lev]; j++)
			for(i = 0; i < o->nox[lev]; i++)
				for(k = 0; k < o->noz[lev]; k++)
				{
					const int n = nidx(o, lev, k, i, j);
					o->X[lev][0][n] = xc[i * s];
					o->X[lev][1][n] = yc[j * s];
					o->X[lev][2][n] = zc[k * s];
				}
	}
	for(d = 0; d < 3; d++)
		o->TB[d] = calloc(o->nno[o->levmax], sizeof(float));
	boundary_flags(o);
	for(lev = o->levmax; lev >= o->levmin; lev--)
	{
		for(e = 0; e < o->nel[lev]; e++)
			global_shape_fn(o, lev, e);
		mass_matrix(o, lev);
	}
	/* material groups, construct_mat_group Construct_arrays.c:742-829 with
	 * layers() Viscosity_structures.c:1389-1404 */
	lev = o->levmax;
	o->mat = malloc(sizeof(int) * o->nel[lev]);
	for(e = 0; e < o->nel[lev]; e++)
	{
		int ez, ex, ey, a, l = 0;
		float x3 = 0.0f;
		edecomp(o, lev, e, &ez, &ex, &ey);
		for(a = 1; a <= 8; a++)
			x3 += o->X[lev][2][enode(o, lev, ez, ex, ey, a)];
		x3 = x3 / 8;
		for(i = 0; i < p->num_mat - 1; i++)
			if(x3 >= p->z_layer[i])
			{
				l = i + 1;
				break;
			}
		o->mat[e] = l ? l : p->num_mat;
	}
	o->expansivity = malloc(sizeof(float) * o->noz[lev]);
	o->diffusivity = malloc(sizeof(float) * o->noz[lev]);
	for(k = 0; k < o->noz[lev]; k++)
	{
		/* thermexp_factor = thermdiff_factor = 1 (Instructions.c:1062-1084) */
		o->expansivity[k] = 1.0f;
		o->diffusivity[k] = 1.0f;
	}
	/* isoviscous default until a viscosity is set */
	for(lev = o->levmin; lev <= o->levmax; lev++)
		for(i = 0; i < o->nel[lev] * 8; i++)
			o->EVI[lev][i] = 1.0f;
	return o;
}

void orc_free(orc *o)
{
	int lev, d;
	if(!o)
		return;
	for(lev = o->levmin; lev <= o->levmax; lev++)
	{
		for(d = 0; d < 3; d++)
			free(o->X[lev][d]);
		free(o->NODE[lev]);
		free(o->GNX[lev]);
		free(o->GDA[lev]);
		free(o->ECO[lev]);
		free(o->TWW[lev]);
		free(o->MASS[lev]);
		free(o->EVI[lev]);
		free(o->EG[lev]);
		free(o->K[lev]);
		free(o->BI[lev]);
		free(o->BPI[lev]);
		free(o->mg_vel[lev]);
		free(o->mg_res[lev]);
		free(o->mg_rhs[lev]);
		free(o->mg_fl[lev]);
		free(o->mg_del[lev]);
		free(o->mg_AU[lev]);
	}
	for(d = 0; d < 3; d++)
		free(o->TB[d]);
	free(o->Mc);
	free(o->Ma);
	free(o->Mb);
	free(o->mat);
	free(o->expansivity);
	free(o->diffusivity);
	free(o);
}

void *orc_get(orc *o, const char *name, int lev, long *count)
{
	long n = 0;
	void *r = NULL;
#define G(nm, ptr, cnt) else if(strcmp(name, nm) == 0) { r = (void *)(ptr); n = (cnt); }
	if(0)
	{
	}
	G("X", o->X[lev][0], o->nno[lev])
	G("Y", o->X[lev][1], o->nno[lev])
	G("Z", o->X[lev][2], o->nno[lev])
	G("NODE", o->NODE[lev], o->nno[lev])
	G("GNX", o->GNX[lev], (long)o->nel[lev] * GNX_SZ)
	G("GDA", o->GDA[lev], (long)o->nel[lev] * GDA_SZ)
	G("ECO", o->ECO[lev], (long)o->nel[lev] * ECO_SZ)
	G("TWW", o->TWW[lev], (long)o->nel[lev] * TWW_SZ)
	G("MASS", o->MASS[lev], o->nno[lev])
	G("EVI", o->EVI[lev], (long)o->nel[lev] * 8)
	G("EG", o->EG[lev], (long)o->nel[lev] * EG_SZ)
	G("K", o->K[lev], (long)o->nno[lev] * KROW)
	G("BI", o->BI[lev], o->neq[lev])
	G("BPI", o->BPI[lev], o->nel[lev])
	G("TB", o->TB[2], o->nno[o->levmax])
	G("mat", o->mat, o->nel[o->levmax])
	G("dims", o->nox, 0)
#undef G
	if(count)
		*count = n;
	return r;
}

/* ------------------------------------------------------------------------ */
/* viscosity                                                                 */

/* visc_from_T, Viscosity_structures.c:491-737 (rheol 0, 3 and 11), followed by
 * the MAX/MIN clip of get_system_viscosity :409-425 */
void orc_visc_from_T(orc *o, const float *T, float *evi)
{
	const int lev = o->levmax;
	int e, jj, kk;
	const float zero = 0.0f, one = 1.0f;
	for(e = 0; e < o->nel[lev]; e++)
	{
		int ez, ex, ey;
		const int l = o->mat[e] - 1;
		const float tempa = o->p.visc0[l];
		float TT[9];
		edecomp(o, lev, e, &ez, &ex, &ey);
		for(kk = 1; kk <= 8; kk++)
			TT[kk] = T[enode(o, lev, ez, ex, ey, kk)];
		for(jj = 1; jj <= 8; jj++)
		{
			float temp = 1.0e-32f, arg;
			double v;
			for(kk = 1; kk <= 8; kk++)
				temp = (float)(temp + fmaxf_(zero, TT[kk]) * o->Nv[GNV(kk, jj)]);
			if(!o->p.tdepv)
				v = tempa;
			else if(o->p.rheol == 11)	/* :697-719 */
				v = tempa * exp(o->p.viscE[l] / (temp + o->p.viscT[l]) - o->p.viscE[l] / (0.5 + o->p.viscT[l]));
			else
			{
				if(o->p.rheol == 3)
					arg = o->p.viscE[l] * (o->p.viscT[l] - temp);
				else
					arg = o->p.viscE[l] * (one - temp);
				v = tempa * exp(arg);
			}
			evi[e * 8 + jj - 1] = (float)v;
		}
	}
	if(o->p.vmax_on)
		for(e = 0; e < o->nel[lev] * 8; e++)
			if(evi[e] > o->p.visc_max)
				evi[e] = o->p.visc_max;
	if(o->p.vmin_on)
		for(e = 0; e < o->nel[lev] * 8; e++)
			if(evi[e] < o->p.visc_min)
				evi[e] = o->p.visc_min;
}

/* project_viscosity with visc_smooth_cycles = 1, Solver_multigrid.c:398-480:
 * visc_from_gint_to_nodes (Nodal_mesh.c:583-615), project_scalar
 * (Solver_multigrid.c:344-388), visc_from_nodes_to_gint (Nodal_mesh.c:617-640) */
static void project_viscosity(orc *o)
{
	int lv;
	for(lv = o->levmax; lv > o->levmin; lv--)
	{
		const int sl = lv - 1;
		float *VN = calloc(o->nno[lv], sizeof(float));
		float *VD = calloc(o->nno[sl], sizeof(float));
		int e, i, j, a;
		for(e = 0; e < o->nel[lv]; e++)
		{
			int ez, ex, ey;
			double tv = 0.0;
			edecomp(o, lv, e, &ez, &ex, &ey);
			for(i = 0; i < 8; i++)
				tv += o->EVI[lv][e * 8 + i];
			tv = tv / 8;
			for(j = 1; j <= 8; j++)
			{
				const int n = enode(o, lv, ez, ex, ey, j);
				VN[n] = (float)(VN[n] + o->TWW[lv][(size_t)e * TWW_SZ + j] * tv);
			}
		}
		for(i = 0; i < o->nno[lv]; i++)
			VN[i] *= o->MASS[lv][i];
		/* project_scalar */
		for(e = 0; e < o->nel[sl]; e++)
		{
			int ez, ex, ey;
			edecomp(o, sl, e, &ez, &ex, &ey);
			for(a = 1; a <= 8; a++)
			{
				/* fine sub-element at corner a (construct_sub_element,
				 * Construct_arrays.c:636-665) */
				const int fz = 2 * ez + offz[a], fx = 2 * ex + offx[a], fy = 2 * ey + offy[a];
				float average = 0.0f, w;
				for(j = 1; j <= 8; j++)
					average += VN[nidx(o, lv, fz + offz[j], fx + offx[j], fy + offy[j])];
				w = (float)((1.0 / 8) * average);
				{
					const int n = enode(o, sl, ez, ex, ey, a);
					VD[n] += w * o->TWW[sl][(size_t)e * TWW_SZ + a];
				}
			}
		}
		for(i = 0; i < o->nno[sl]; i++)
			VD[i] *= o->MASS[sl][i];
		for(e = 0; e < o->nel[sl]; e++)
		{
			int ez, ex, ey;
			edecomp(o, sl, e, &ez, &ex, &ey);
			for(i = 1; i <= 8; i++)
			{
				double tv = 0.0;
				for(j = 1; j <= 8; j++)
					tv += o->Nv[GNV(j, i)] * VD[enode(o, sl, ez, ex, ey, j)];
				o->EVI[sl][e * 8 + i - 1] = (float)tv;
			}
		}
		free(VN);
		free(VD);
	}
}

/* ------------------------------------------------------------------------ */
/* element matrices                                                          */

/* get_elt_k, Element_calculations.c:133-293 (cartesian isotropic branch);
 * fills the full 24x24 elt_k exactly as the reference does */
static void elt_k(orc *o, int lev, int el, double K[24 * 24])
{
	const int n = 24;
	double W[9];
	const float *gnx = o->GNX[lev] + (size_t)el * GNX_SZ;
	const float *gda = o->GDA[lev] + (size_t)el * GDA_SZ;
	int a, b, k;
	for(k = 1; k <= 8; k++)
	{
		float w = 1.0f * gda[k];
		w = w * o->EVI[lev][el * 8 + k - 1];
		W[k] = w;
	}
	for(a = 1; a <= 8; a++)
		for(b = a; b <= 8; b++)
		{
			double m[4][4] = { {0} }, temp = 0.0;
			int pn, qn, ad, bd;
			for(k = 1; k <= 8; k++)
			{
				m[1][1] += W[k] * gnx[GNVX(0, a, k)] * gnx[GNVX(0, b, k)];
				m[1][2] += W[k] * gnx[GNVX(1, a, k)] * gnx[GNVX(0, b, k)];
				m[1][3] += W[k] * gnx[GNVX(2, a, k)] * gnx[GNVX(0, b, k)];
				m[2][1] += W[k] * gnx[GNVX(0, a, k)] * gnx[GNVX(1, b, k)];
				m[2][2] += W[k] * gnx[GNVX(1, a, k)] * gnx[GNVX(1, b, k)];
				m[2][3] += W[k] * gnx[GNVX(2, a, k)] * gnx[GNVX(1, b, k)];
				m[3][1] += W[k] * gnx[GNVX(0, a, k)] * gnx[GNVX(2, b, k)];
				m[3][2] += W[k] * gnx[GNVX(1, a, k)] * gnx[GNVX(2, b, k)];
				m[3][3] += W[k] * gnx[GNVX(2, a, k)] * gnx[GNVX(2, b, k)];
			}
			for(k = 1; k <= 8; k++)
			{
				/* float products summed in float, then scaled in double */
				const float s = gnx[GNVX(0, a, k)] * gnx[GNVX(0, b, k)] + gnx[GNVX(1, a, k)] * gnx[GNVX(1, b, k)] + gnx[GNVX(2, a, k)] * gnx[GNVX(2, b, k)];
				temp += W[k] * s;
			}
			m[1][1] += temp;
			m[2][2] += temp;
			m[3][3] += temp;
			ad = 3 * (a - 1);
			bd = 3 * (b - 1);
			pn = ad * n + bd;
			qn = bd * n + ad;
			K[pn] = m[1][1];
			K[pn + 1] = m[1][2];
			K[pn + 2] = m[1][3];
			K[pn + n] = m[2][1];
			K[pn + n + 1] = m[2][2];
			K[pn + n + 2] = m[2][3];
			K[pn + 2 * n] = m[3][1];
			K[pn + 2 * n + 1] = m[3][2];
			K[pn + 2 * n + 2] = m[3][3];
			K[qn] = m[1][1];
			K[qn + 1] = m[2][1];
			K[qn + 2] = m[3][1];
			K[qn + n] = m[1][2];
			K[qn + n + 1] = m[2][2];
			K[qn + n + 2] = m[3][2];
			K[qn + 2 * n] = m[1][3];
			K[qn + 2 * n + 1] = m[2][3];
			K[qn + 2 * n + 2] = m[3][3];
		}
}

/* get_elt_g, Element_calculations.c:831-934 (CART3D branch) */
static void elt_g(orc *o, int lev, int el)
{
	const float *gnx = o->GNX[lev] + (size_t)el * GNX_SZ;
	const float *gda = o->GDA[lev] + (size_t)el * GDA_SZ;
	float *g = o->EG[lev] + (size_t)el * EG_SZ;
	const double temp = 8.0f * gda[10];	/* p_point[1].weight[2] * GDA.ppt[1] */
	int a;
	for(a = 1; a <= 8; a++)
	{
		const int p = 3 * (a - 1);
		g[p] = (float)(-gnx[192 + GNPX(0, a)] * temp);
		g[p + 1] = (float)(-gnx[192 + GNPX(1, a)] * temp);
		g[p + 2] = (float)(-gnx[192 + GNPX(2, a)] * temp);
	}
}

/* get_aug_k, Element_calculations.c:1127-1163 */
static void aug_k(orc *o, int lev, int el, double K[24 * 24])
{
	const int n = 24;
	const float *g = o->EG[lev] + (size_t)el * EG_SZ;
	double Visc = 0.0;
	int a, b, r, c;
	for(a = 0; a < 8; a++)
		Visc += o->EVI[lev][el * 8 + a];
	Visc = Visc / 8;
	for(a = 1; a <= 8; a++)
		for(b = 1; b <= 8; b++)
		{
			const int i = (a - 1) * n * 3 + (b - 1) * 3;
			/* the reference's 9 statements, in their order, reduce to this */
			for(r = 0; r < 3; r++)
				for(c = 0; c < 3; c++)
					K[i + r * n + c] += Visc * o->p.aug_number * g[3 * (a - 1) + r] * g[3 * (b - 1) + c];
		}
}

/* neighbour slot of node (dz,dx,dy) offsets in the 81-wide row */
static inline int nslot(int dz, int dx, int dy)
{
	return (dz + 1) + 3 * (dx + 1) + 9 * (dy + 1);
}

/* construct_node_ks, Construct_arrays.c:359-535: the reference stores, for
 * each node, the 3x3 blocks coupling it to itself and to its 13 lower-
 * numbered neighbours (symmetric half, float, accumulated element by element
 * in element order with BC weights).  The restatement keeps exactly those
 * accumulations and expands them to the full 27-neighbour row:
 *   K[n][slot(m)][r][c] = stored_{max(n,m)}(min(n,m)) (transposed if m > n).
 * Also build_diagonal_of_K (:624-652) into BI. */
static void node_ks(orc *o, int lev)
{
	const int nno = o->nno[lev];
	float *half = calloc((size_t)nno * KROW, sizeof(float));	/* [n][slot][r][c], slot of the lower node */
	double *diag = o->BI[lev];
	double K[24 * 24];
	int e, i, j, r, c, n;
	for(i = 0; i < o->neq[lev]; i++)
		diag[i] = 0.0;
	for(e = 0; e < o->nel[lev]; e++)
	{
		int ez, ex, ey, nodes[9];
		edecomp(o, lev, e, &ez, &ex, &ey);
		for(i = 1; i <= 8; i++)
			nodes[i] = enode(o, lev, ez, ex, ey, i);
		elt_k(o, lev, e, K);
		if(o->p.aug_lagr)
			aug_k(o, lev, e, K);
		/* build_diagonal_of_K */
		for(i = 1; i <= 8; i++)
			for(r = 0; r < 3; r++)
			{
				const int pp = 3 * (i - 1) + r;
				diag[3 * nodes[i] + r] += K[pp * 24 + pp];
			}
		for(i = 1; i <= 8; i++)
		{
			const int node = nodes[i];
			const unsigned fa = o->NODE[lev][node];
			const double w[3] = { (fa & ORC_VBX) ? 0.0 : 1.0, (fa & ORC_VBY) ? 0.0 : 1.0, (fa & ORC_VBZ) ? 0.0 : 1.0 };
			for(j = 1; j <= 8; j++)
			{
				const int node1 = nodes[j];
				if(node1 <= node)
				{
					const unsigned fb = o->NODE[lev][node1];
					const double ww[3] = { (fb & ORC_VBX) ? 0.0 : 1.0, (fb & ORC_VBY) ? 0.0 : 1.0, (fb & ORC_VBZ) ? 0.0 : 1.0 };
					const int s = nslot(offz[j] - offz[i], offx[j] - offx[i], offy[j] - offy[i]);
					const int pp = 3 * (i - 1), qq = 3 * (j - 1);
					/* reference order: column direction outer, row inner */
					for(c = 0; c < 3; c++)
						for(r = 0; r < 3; r++)
						{
							float *dst = &half[(size_t)node * KROW + s * 9 + r * 3 + c];
							*dst = (float)(*dst + w[r] * ww[c] * K[(pp + r) * 24 + qq + c]);
						}
				}
			}
		}
	}
	/* expand to the full row */
	for(n = 0; n < nno; n++)
	{
		float *row = o->K[lev] + (size_t)n * KROW;
		const int k = n % o->noz[lev], i2 = (n / o->noz[lev]) % o->nox[lev], j2 = n / (o->noz[lev] * o->nox[lev]);
		int dz, dx, dy;
		for(i = 0; i < KROW; i++)
			row[i] = 0.0f;
		for(dy = -1; dy <= 1; dy++)
			for(dx = -1; dx <= 1; dx++)
				for(dz = -1; dz <= 1; dz++)
				{
					const int kk = k + dz, ii = i2 + dx, jj = j2 + dy;
					int m, s;
					if(kk < 0 || kk >= o->noz[lev] || ii < 0 || ii >= o->nox[lev] || jj < 0 || jj >= o->noy[lev])
						continue;
					m = nidx(o, lev, kk, ii, jj);
					s = nslot(dz, dx, dy);
					for(r = 0; r < 3; r++)
						for(c = 0; c < 3; c++)
						{
							float v;
							/* a node's own block (m == n) is read transposed by the
							 * reference's matvec and smoother (Element_calculations.c:604-605,
							 * General_matrix_functions.c:1254-1255) */
							if(m < n)
								v = half[(size_t)n * KROW + s * 9 + r * 3 + c];
							else
								v = half[(size_t)m * KROW + nslot(-dz, -dx, -dy) * 9 + c * 3 + r];
							row[r * 81 + s * 3 + c] = v;
						}
				}
	}
	for(i = 0; i < o->neq[lev]; i++)
		diag[i] = (float)1.0 / diag[i];
	free(half);
}

/* assemble_dAhatp_entry / build_diagonal_of_Ahat, Element_calculations.c:654-685, 771-824 */
static void build_bpi(orc *o, int lev)
{
	int e, a, d;
	for(e = 0; e < o->nel[lev]; e++)
	{
		int ez, ex, ey;
		const float *g = o->EG[lev] + (size_t)e * EG_SZ;
		double gradP[24], divU = 0.0;
		o->BPI[lev][e] = 1.0;
		if(!o->p.precond)
			continue;
		edecomp(o, lev, e, &ez, &ex, &ey);
		for(a = 1; a <= 8; a++)
		{
			const int n = enode(o, lev, ez, ex, ey, a);
			for(d = 0; d < 3; d++)
				gradP[3 * (a - 1) + d] = o->BI[lev][3 * n + d] * g[3 * (a - 1) + d];
		}
		for(a = 0; a < 24; a++)
			divU += g[a] * gradP[a];
		o->BPI[lev][e] = (divU != 0.0) ? 1.0 / divU : 1.0;
	}
}

/* construct_stiffness_B_matrix, Construct_arrays.c:834-889 */
/* Smoother calls applied as matrices (8-colour smoother only), on levels of
 * at most MATRIX_MAX_DOF unknowns.
 * The coarsest level's gauss_seidel call in multi_grid (v_steps_low sweeps
 * from a zero guess, General_matrix_functions.c:574, :613) is linear in its
 * right-hand side apart from the FP32 rounding of each increment: column j
 * of Mc is the smoother's result for the unit right-hand side e_j.  The
 * level above (down_heavy = up_heavy sweeps, with or without a guess) is
 * affine: d = Ma din + Mb F, column j of Ma from din = e_j, F = 0, of Mb from
 * din = 0, F = e_j.  Built whenever the operators are rebuilt; they equal
 * the sweeps up to the FP32 rounding of the increments. */
#define MATRIX_MAX_DOF 1500

static void build_affine_operator(orc *o)
{
	const int lev = o->levmin + 1, n3 = o->neq[lev];
	double *e = calloc(n3, sizeof(double)), *z = calloc(n3, sizeof(double)), *Ad = malloc(sizeof(double) * n3);
	int j;
	free(o->Ma);
	free(o->Mb);
	o->Ma = o->Mb = NULL;
	if(!(o->Mc && lev < o->levmax && o->p.down_heavy == o->p.up_heavy && n3 <= MATRIX_MAX_DOF))
		goto done;
	o->Ma = malloc(sizeof(double) * (size_t)n3 * n3);
	o->Mb = malloc(sizeof(double) * (size_t)n3 * n3);
	for(j = 0; j < n3; j++)
	{
		double *ca = o->Ma + (size_t)j * n3, *cb = o->Mb + (size_t)j * n3;
		memset(ca, 0, sizeof(double) * n3);
		ca[j] = 1.0;
		orc_gauss_seidel(o, lev, ca, z, Ad, o->p.down_heavy, 1);
		e[j] = 1.0;
		orc_gauss_seidel(o, lev, cb, e, Ad, o->p.down_heavy, 0);
		e[j] = 0.0;
	}
done:
	free(e);
	free(z);
	free(Ad);
}

/* out = Ma din + Mb F (din NULL: zero guess); row i sums the Ma terms
 * j = t, t+64, ... in order, then the Mb terms likewise, into S[t] (t < 64),
 * and the S[t] meet in a pairwise tree (S[q] += S[q+w], w = 1 .. 32) */
void orc_affine_solve(orc *o, const double *din, const double *F, double *out)
{
	const int n3 = o->neq[o->levmin + 1];
	int i, j, w, q;
	for(i = 0; i < n3; i++)
	{
		double S[64];
		for(q = 0; q < 64; q++)
			S[q] = 0.0;
		if(din)
			for(j = 0; j < n3; j++)
				S[j & 63] = S[j & 63] + o->Ma[(size_t)j * n3 + i] * din[j];
		for(j = 0; j < n3; j++)
			S[j & 63] = S[j & 63] + o->Mb[(size_t)j * n3 + i] * F[j];
		for(w = 1; w < 64; w *= 2)
			for(q = 0; q < 64; q += 2 * w)
				S[q] = S[q] + S[q + w];
		out[i] = S[0];
	}
}

static void build_coarse_operator(orc *o)
{
	const int lev = o->levmin, n3 = o->neq[lev];
	double *e = calloc(n3, sizeof(double)), *Ad = malloc(sizeof(double) * n3);
	int j;
	free(o->Mc);
	o->Mc = NULL;
	if(n3 > MATRIX_MAX_DOF)
	{
		free(e);
		free(Ad);
		return;
	}
	o->Mc = malloc(sizeof(double) * (size_t)n3 * n3);
	for(j = 0; j < n3; j++)
	{
		e[j] = 1.0;
		orc_gauss_seidel(o, lev, o->Mc + (size_t)j * n3, e, Ad, o->p.v_steps_low, 0);
		e[j] = 0.0;
	}
	free(e);
	free(Ad);
}

/* d = Mc F; row i sums j = t, t+16, t+32, ... into S[t] (t < 16), then the
 * S[t] meet in a pairwise tree: S[q] += S[q+w] for w = 1, 2, 4, 8 and q a
 * multiple of 2w; d[i] = S[0] */
void orc_coarse_solve(orc *o, const double *F, double *d)
{
	const int n3 = o->neq[o->levmin];
	int i, j, w, q;
	for(i = 0; i < n3; i++)
	{
		double S[16];
		for(q = 0; q < 16; q++)
			S[q] = 0.0;
		for(j = 0; j < n3; j++)
			S[j & 15] = S[j & 15] + o->Mc[(size_t)j * n3 + i] * F[j];
		for(w = 1; w < 16; w *= 2)
			for(q = 0; q < 16; q += 2 * w)
				S[q] = S[q] + S[q + w];
		d[i] = S[0];
	}
}

void orc_build_operators(orc *o)
{
	int lev, e;
	project_viscosity(o);
	for(lev = o->levmin; lev <= o->levmax; lev++)
		for(e = 0; e < o->nel[lev]; e++)
			elt_g(o, lev, e);
	for(lev = o->levmax; lev >= o->levmin; lev--)
		node_ks(o, lev);
	for(lev = o->levmax; lev >= o->levmin; lev--)
		build_bpi(o, lev);
	/* rebuild_BI_on_boundary (:892-952) changes only OFFSIDE nodes, which a
	 * single subdomain does not have */
	if(o->p.smoother == 1)
	{
		build_coarse_operator(o);
		build_affine_operator(o);
	}
}

void orc_set_viscosity(orc *o, const float *evi)
{
	memcpy(o->EVI[o->levmax], evi, sizeof(float) * 8 * o->nel[o->levmax]);
	orc_build_operators(o);
}

/* ------------------------------------------------------------------------ */
/* vector operators                                                          */

/* strip_bcs_from_residual, Boundary_conditions.c:926-947 */
void orc_strip(orc *o, int lev, double *v)
{
	int n;
	for(n = 0; n < o->nno[lev]; n++)
	{
		const unsigned f = o->NODE[lev][n];
		if(f & ORC_VBX)
			v[3 * n] = 0.0;
		if(f & ORC_VBY)
			v[3 * n + 1] = 0.0;
		if(f & ORC_VBZ)
			v[3 * n + 2] = 0.0;
	}
}

/* row product of node n with vector u (full 81-wide row) */
static inline void row_apply(const orc *o, int lev, int n, const double *u, double out[3])
{
	const float *row = o->K[lev] + (size_t)n * KROW;
	const int noz = o->noz[lev], nox = o->nox[lev], noy = o->noy[lev];
	const int k = n % noz, i = (n / noz) % nox, j = n / (noz * nox);
	int dz, dx, dy, r;
	out[0] = out[1] = out[2] = 0.0;
	for(dy = -1; dy <= 1; dy++)
		for(dx = -1; dx <= 1; dx++)
			for(dz = -1; dz <= 1; dz++)
			{
				const int kk = k + dz, ii = i + dx, jj = j + dy;
				int m, s;
				if(kk < 0 || kk >= noz || ii < 0 || ii >= nox || jj < 0 || jj >= noy)
					continue;
				m = nidx(o, lev, kk, ii, jj);
				s = nslot(dz, dx, dy);
				for(r = 0; r < 3; r++)
					out[r] += row[r * 81 + s * 3 + 0] * u[3 * m] + row[r * 81 + s * 3 + 1] * u[3 * m + 1] + row[r * 81 + s * 3 + 2] * u[3 * m + 2];
			}
}

/* n_assemble_del2_u, Element_calculations.c:552-621 (as a full-row gather) */
void orc_matvec(orc *o, int lev, const double *u, double *Au, int strip)
{
	int n;
	for(n = 0; n < o->nno[lev]; n++)
		row_apply(o, lev, n, u, Au + 3 * n);
	if(strip)
		orc_strip(o, lev, Au);
}

/* The same row product summed in the 8-colour smoother's fixed order: the
 * block terms t[p][b] = k0*u0 + k1*u1 + k2*u2 of each y-plane p = dy+1
 * (b = (dz+1) + 3(dx+1), absent neighbours and b >= 9 are 0) are summed in a
 * pairwise tree over 16 slots, t[q] += t[q+w] for w = 1, 2, 4, 8 and q a
 * multiple of 2w, giving P[p] = t[p][0]; the row sum is (P[0] + P[1]) + P[2].
 * (On the GPU one wavefront handles a plane, or a 16-lane row of lanes
 * does, which is why the planes are summed separately.)  The matvec keeps
 * the sequential order of row_apply. */
static inline void row_apply_tree(const orc *o, int lev, int n, const double *u, double out[3])
{
	const float *row = o->K[lev] + (size_t)n * KROW;
	const int noz = o->noz[lev], nox = o->nox[lev], noy = o->noy[lev];
	const int k = n % noz, i = (n / noz) % nox, j = n / (noz * nox);
	double P[3][3];
	int p, b, r, w, q;
	for(p = 0; p < 3; p++)
	{
		double t[3][16];
		for(r = 0; r < 3; r++)
			for(b = 0; b < 16; b++)
				t[r][b] = 0.0;
		for(b = 0; b < 9; b++)
		{
			const int dz = b % 3 - 1, dx = b / 3 - 1, dy = p - 1;
			const int kk = k + dz, ii = i + dx, jj = j + dy;
			const int s = nslot(dz, dx, dy);
			int m;
			if(kk < 0 || kk >= noz || ii < 0 || ii >= nox || jj < 0 || jj >= noy)
				continue;
			m = nidx(o, lev, kk, ii, jj);
			for(r = 0; r < 3; r++)
				t[r][b] = row[r * 81 + s * 3 + 0] * u[3 * m] + row[r * 81 + s * 3 + 1] * u[3 * m + 1] + row[r * 81 + s * 3 + 2] * u[3 * m + 2];
		}
		for(w = 1; w < 16; w *= 2)
			for(q = 0; q < 16; q += 2 * w)
				for(r = 0; r < 3; r++)
					t[r][q] = t[r][q] + t[r][q + w];
		for(r = 0; r < 3; r++)
			P[r][p] = t[r][0];
	}
	for(r = 0; r < 3; r++)
		out[r] = (P[r][0] + P[r][1]) + P[r][2];
}

static void node_update(orc *o, int lev, int n, double *d0, const double *F)
{
	double Ad[3];
	float t[3];
	int r;
	if(o->p.smoother == 0)
		row_apply(o, lev, n, d0, Ad);
	else
		row_apply_tree(o, lev, n, d0, Ad);
	/* temp is higher_precision (float) in the reference (:1169, :1236-1238) */
	for(r = 0; r < 3; r++)
		t[r] = (float)((F[3 * n + r] - Ad[r]) * o->BI[lev][3 * n + r]);
	for(r = 0; r < 3; r++)
		d0[3 * n + r] += t[r];
}

/* gauss_seidel, General_matrix_functions.c:1160-1361.
 * smoother 0: the reference's lexicographic node-block sweep.  Its
 * incremental Ad update is restated as a fresh row product per node, which
 * gives the same iterate up to the float rounding of the reference's
 * scattered Ad increments.
 * smoother 1: the same node-block update in 8 colour classes of the node
 * index parities (k%2, i%2, j%2), colour c = (k&1) | (i&1)<<1 | (j&1)<<2,
 * with the row sums of row_apply_tree.
 * Both return Ad = K d0 (BC rows stripped). */
void orc_gauss_seidel(orc *o, int lev, double *d0, const double *F, double *Ad, int cycles, int guess)
{
	const int nno = o->nno[lev];
	int c, n, cyc;
	if(!guess)
		for(n = 0; n < 3 * nno; n++)
			d0[n] = 0.0;
	for(cyc = 0; cyc < cycles; cyc++)
	{
		if(o->p.smoother == 0)
		{
			for(n = 0; n < nno; n++)
				node_update(o, lev, n, d0, F);
		}
		else
		{
			for(c = 0; c < 8; c++)
				for(n = 0; n < nno; n++)
				{
					const int k = n % o->noz[lev], i = (n / o->noz[lev]) % o->nox[lev], j = n / (o->noz[lev] * o->nox[lev]);
					if(((k & 1) | ((i & 1) << 1) | ((j & 1) << 2)) == c)
						node_update(o, lev, n, d0, F);
				}
		}
	}
	orc_matvec(o, lev, d0, Ad, 1);
}

/* project_vector, Solver_multigrid.c:72-159 (restriction lev -> lev-1) */
void orc_project(orc *o, int lev, const double *AU, double *AD)
{
	const int sl = lev - 1;
	int e, a, j, d, n;
	for(n = 0; n < o->neq[sl]; n++)
		AD[n] = 0.0;
	for(e = 0; e < o->nel[sl]; e++)
	{
		int ez, ex, ey;
		edecomp(o, sl, e, &ez, &ex, &ey);
		for(a = 1; a <= 8; a++)
		{
			const int fz = 2 * ez + offz[a], fx = 2 * ex + offx[a], fy = 2 * ey + offy[a];
			double av[3] = { 0, 0, 0 };
			const int cn = enode(o, sl, ez, ex, ey, a);
			for(j = 1; j <= 8; j++)
			{
				const int fn = nidx(o, lev, fz + offz[j], fx + offx[j], fy + offy[j]);
				for(d = 0; d < 3; d++)
					av[d] += AU[3 * fn + d];
			}
			for(d = 0; d < 3; d++)
				AD[3 * cn + d] += o->TWW[sl][(size_t)e * TWW_SZ + a] * av[d];
		}
	}
	for(n = 0; n < o->nno[sl]; n++)
		for(d = 0; d < 3; d++)
			AD[3 * n + d] = AD[3 * n + d] * o->MASS[sl][n];
}

/* interp_vector + un_inject_vector, Solver_multigrid.c:173-298, 581-634
 * (prolongation lev-1 -> lev) */
void orc_interp(orc *o, int lev, const double *AD, double *AU)
{
	const int cl = lev - 1;
	const int nox = o->nox[lev], noz = o->noz[lev], noy = o->noy[lev];
	const float *eco = o->ECO[lev];
	int i, j, k, d;
	/* the first element (lowest number) that holds a node: NEI[lev].element */
#define FIRST_EL(kk, ii, jj) (((kk) > 0 ? (kk) - 1 : 0) + o->elz[lev] * (((ii) > 0 ? (ii) - 1 : 0) + o->elx[lev] * ((jj) > 0 ? (jj) - 1 : 0)))
	for(j = 0; j < o->noy[cl]; j++)
		for(i = 0; i < o->nox[cl]; i++)
			for(k = 0; k < o->noz[cl]; k++)
				for(d = 0; d < 3; d++)
					AU[3 * nidx(o, lev, 2 * k, 2 * i, 2 * j) + d] = AD[3 * nidx(o, cl, k, i, j) + d];
	for(j = 0; j < noy; j += 2)	/* x gaps */
		for(k = 0; k < noz; k += 2)
			for(i = 1; i < nox - 1; i += 2)
			{
				const float x1 = eco[(size_t)FIRST_EL(k, i - 1, j) * ECO_SZ + 4 + 1];
				const float x2 = eco[(size_t)FIRST_EL(k, i + 1, j) * ECO_SZ + 4 + 1];
				const float n1 = x2 / (x1 + x2), n2 = x1 / (x1 + x2);
				const int n0 = nidx(o, lev, k, i, j), m1 = nidx(o, lev, k, i - 1, j), m2 = nidx(o, lev, k, i + 1, j);
				for(d = 0; d < 3; d++)
					AU[3 * n0 + d] = n1 * AU[3 * m1 + d] + n2 * AU[3 * m2 + d];
			}
	for(j = 0; j < noy; j += 2)	/* z gaps */
		for(i = 0; i < nox; i++)
			for(k = 1; k < noz - 1; k += 2)
			{
				const float x1 = eco[(size_t)FIRST_EL(k - 1, i, j) * ECO_SZ + 4 + 3];
				const float x2 = eco[(size_t)FIRST_EL(k + 1, i, j) * ECO_SZ + 4 + 3];
				const float n1 = x2 / (x1 + x2), n2 = x1 / (x1 + x2);
				const int n0 = nidx(o, lev, k, i, j), m1 = nidx(o, lev, k - 1, i, j), m2 = nidx(o, lev, k + 1, i, j);
				for(d = 0; d < 3; d++)
					AU[3 * n0 + d] = n1 * AU[3 * m1 + d] + n2 * AU[3 * m2 + d];
			}
	for(i = 0; i < nox; i++)	/* y gaps */
		for(k = 0; k < noz; k++)
			for(j = 1; j < noy - 1; j += 2)
			{
				const float x1 = eco[(size_t)FIRST_EL(k, i, j - 1) * ECO_SZ + 4 + 2];
				const float x2 = eco[(size_t)FIRST_EL(k, i, j + 1) * ECO_SZ + 4 + 2];
				const float n1 = x2 / (x1 + x2), n2 = x1 / (x1 + x2);
				const int n0 = nidx(o, lev, k, i, j), m1 = nidx(o, lev, k, i, j - 1), m2 = nidx(o, lev, k, i, j + 1);
				for(d = 0; d < 3; d++)
					AU[3 * n0 + d] = n1 * AU[3 * m1 + d] + n2 * AU[3 * m2 + d];
			}
#undef FIRST_EL
}

/* assemble_div_u, Element_calculations.c:691-720 */
void orc_div(orc *o, int lev, const double *U, double *divU)
{
	int e, a;
	for(e = 0; e < o->nel[lev]; e++)
		divU[e] = 0.0;
	for(a = 1; a <= 8; a++)
		for(e = 0; e < o->nel[lev]; e++)
		{
			int ez, ex, ey;
			const float *g = o->EG[lev] + (size_t)e * EG_SZ;
			const int p = 3 * (a - 1);
			int n;
			edecomp(o, lev, e, &ez, &ex, &ey);
			n = enode(o, lev, ez, ex, ey, a);
			divU[e] += g[p] * U[3 * n] + g[p + 1] * U[3 * n + 1] + g[p + 2] * U[3 * n + 2];
		}
}

/* assemble_grad_p, Element_calculations.c:727-769 */
void orc_grad(orc *o, int lev, const double *P, double *gradP)
{
	int e, a, i;
	for(i = 0; i < o->neq[lev]; i++)
		gradP[i] = 0.0;
	for(e = 0; e < o->nel[lev]; e++)
	{
		int ez, ex, ey;
		const float *g = o->EG[lev] + (size_t)e * EG_SZ;
		if(0.0 == P[e])
			continue;
		edecomp(o, lev, e, &ez, &ex, &ey);
		for(a = 1; a <= 8; a++)
		{
			const int p = 3 * (a - 1), n = enode(o, lev, ez, ex, ey, a);
			gradP[3 * n] += g[p] * P[e];
			gradP[3 * n + 1] += g[p + 1] * P[e];
			gradP[3 * n + 2] += g[p + 2] * P[e];
		}
	}
	orc_strip(o, lev, gradP);
}

static double vdot(const orc *o, int lev, const double *A, const double *B)
{
	double s = 0.0;
	int i;
	for(i = 0; i < o->neq[lev]; i++)
		s += A[i] * B[i];
	return s;
}

static double pdot(const orc *o, int lev, const double *A, const double *B)
{
	double s = 0.0;
	int i;
	for(i = 0; i < o->nel[lev]; i++)
		s += A[i] * B[i];
	return s;
}

/* ------------------------------------------------------------------------ */
/* multigrid and Uzawa                                                       */

/* the coarsest-level smoother call of multi_grid (AU[levmin] is not read
 * afterwards) */
static void coarse_smooth(orc *o, double *vel, const double *rhs, double *AU)
{
	if(o->p.smoother == 1 && o->Mc)
		orc_coarse_solve(o, rhs, vel);
	else
		orc_gauss_seidel(o, o->levmin, vel, rhs, AU, o->p.v_steps_low, 0);
}

/* a smoother call of multi_grid above the coarsest level: the affine form on
 * the level that has one (Ad = K d by a matvec, as gauss_seidel returns) */
static void mg_smooth(orc *o, int lev, double *d, const double *F, double *Ad, int cycles, int guess)
{
	if(o->Ma && lev == o->levmin + 1 && cycles == o->p.down_heavy)
	{
		const int n3 = o->neq[lev];
		double *out = malloc(sizeof(double) * n3);
		orc_affine_solve(o, guess ? d : NULL, F, out);
		memcpy(d, out, sizeof(double) * n3);
		free(out);
		orc_matvec(o, lev, d, Ad, 1);
	}
	else
		orc_gauss_seidel(o, lev, d, F, Ad, cycles, guess);
}

/* multi_grid, General_matrix_functions.c:525-653 */
double orc_multi_grid(orc *o, double *d1, double *F, double *Au, double acc, int hl)
{
	const int levmin = o->levmin, levmax = hl;
	double **vel = o->mg_vel, **res = o->mg_res, **rhs = o->mg_rhs, **fl = o->mg_fl, **del = o->mg_del, **AU = o->mg_AU;
	int lev, dlev, ulev, i, Vn;
	(void)Au;
	(void)acc;
	for(i = 0; i < o->neq[levmax]; i++)
		fl[levmax][i] = F[i];
	for(lev = levmax; lev > levmin; lev--)
	{
		orc_project(o, lev, fl[lev], fl[lev - 1]);
		orc_strip(o, lev - 1, fl[lev - 1]);
	}
	coarse_smooth(o, vel[levmin], fl[levmin], AU[levmin]);
	for(lev = levmin + 1; lev <= levmax; lev++)
	{
		orc_interp(o, lev, vel[lev - 1], vel[lev]);
		orc_strip(o, lev, vel[lev]);
		for(i = 0; i < o->neq[lev]; i++)
			rhs[lev][i] = fl[lev][i];
		for(Vn = 1; Vn <= o->p.mg_cycle; Vn++)
		{
			for(dlev = lev; dlev >= levmin + 1; dlev--)
			{
				const int cycles = (dlev == levmax) ? o->p.v_steps_high : o->p.down_heavy;
				mg_smooth(o, dlev, vel[dlev], rhs[dlev], AU[dlev], cycles, dlev == lev);
				for(i = 0; i < o->neq[dlev]; i++)
					res[dlev][i] = rhs[dlev][i] - AU[dlev][i];
				orc_project(o, dlev, res[dlev], rhs[dlev - 1]);
				orc_strip(o, dlev - 1, rhs[dlev - 1]);
			}
			coarse_smooth(o, vel[levmin], rhs[levmin], AU[levmin]);
			for(ulev = levmin + 1; ulev <= lev; ulev++)
			{
				const int cycles = (ulev == levmax) ? o->p.v_steps_high : o->p.up_heavy;
				double AudotAu, alpha;
				orc_interp(o, ulev, vel[ulev - 1], del[ulev]);
				orc_strip(o, ulev, del[ulev]);
				mg_smooth(o, ulev, del[ulev], res[ulev], AU[ulev], cycles, 1);
				AudotAu = vdot(o, ulev, AU[ulev], AU[ulev]);
				alpha = vdot(o, ulev, AU[ulev], res[ulev]) / AudotAu;
				for(i = 0; i < o->neq[ulev]; i++)
					vel[ulev][i] += alpha * del[ulev][i];
				if(ulev == levmax)
					for(i = 0; i < o->neq[ulev]; i++)
						res[ulev][i] -= alpha * AU[ulev][i];
			}
		}
	}
	for(i = 0; i < o->neq[levmax]; i++)
	{
		F[i] = res[levmax][i];
		d1[i] = vel[levmax][i];
	}
	return sqrt(vdot(o, levmax, F, F) / o->neq[levmax]);
}

/* solve_del2_u, General_matrix_functions.c:368-520 (multigrid branch) */
int orc_solve_del2_u(orc *o, double *d0, double *F, double acc, int hl, int *count)
{
	const int neq = o->neq[hl];
	double *r = malloc(sizeof(double) * neq), *D1 = malloc(sizeof(double) * neq);
	double residual, r0;
	int i, valid, cnt = 0;
	for(i = 0; i < neq; i++)
	{
		r[i] = F[i];
		d0[i] = D1[i] = 0.0;
	}
	r0 = residual = sqrt(vdot(o, hl, r, r) / neq);
	acc = fmax(acc, r0 * o->p.accuracy);
	valid = (residual < acc) ? 0 : 1;
	while(residual > acc)
	{
		residual = orc_multi_grid(o, D1, r, NULL, acc, hl);
		for(i = 0; i < neq; i++)
		{
			d0[i] += D1[i];
			D1[i] = 0.0;
		}
		cnt++;
	}
	if(count)
		*count = cnt;
	free(r);
	free(D1);
	return valid;
}

/* solve_Ahat_p_fhat, Stokes_flow_Incomp.c:295-497, followed by v_from_vector
 * semantics left to the caller.  stats[0] = iterations, [1] = vdotv,
 * [2] = pdotp, [3] = incompressibility, [4] = total inner MG cycles */
int orc_solve_stokes(orc *o, double *V, double *P, const double *F, double *stats)
{
	const int lev = o->levmax, npno = o->nel[lev], neq = o->neq[lev];
	const double imp = o->p.accuracy;
	double *r0 = calloc(npno, sizeof(double)), *r1 = calloc(npno, sizeof(double)), *r2 = calloc(npno, sizeof(double));
	double *z0 = calloc(npno, sizeof(double)), *z1 = calloc(npno, sizeof(double));
	double *s1 = calloc(npno, sizeof(double)), *s2 = calloc(npno, sizeof(double));
	double *Ah = calloc(neq, sizeof(double)), *u1 = calloc(neq, sizeof(double)), *sh;
	double v_res, r1dotz1, r0dotr0 = 0, delta, s2dotAhat, alpha;
	double vdotv, pdotp, incomp = 0;
	float dpressure = 1.0f, dvelocity = 1.0f;
	int i, count = 0, valid, mgc, mgtot = 0;
	const int gneq = neq, gnpno = npno;

	v_res = sqrt(vdot(o, lev, F, F) / gneq);
	orc_grad(o, lev, P, Ah);
	orc_matvec(o, lev, V, u1, 1);
	for(i = 0; i < neq; i++)
		Ah[i] = F[i] - Ah[i] - u1[i];
	orc_strip(o, lev, Ah);
	valid = orc_solve_del2_u(o, u1, Ah, imp * v_res, lev, &mgc);
	mgtot += mgc;
	orc_strip(o, lev, u1);
	for(i = 0; i < neq; i++)
		V[i] += u1[i];
	orc_div(o, lev, V, r1);
	pdotp = sqrt(pdot(o, lev, r1, r1) / gnpno);
	vdotv = sqrt(vdot(o, lev, V, V) / gneq);
	incomp = pdotp / vdotv;

	while((count < o->p.p_iterations) && (dpressure >= imp || dvelocity >= imp))
	{
		for(i = 0; i < npno; i++)
			z1[i] = o->BPI[lev][i] * r1[i];
		r1dotz1 = pdot(o, lev, r1, z1);
		if(count == 0)
			for(i = 0; i < npno; i++)
				s2[i] = z1[i];
		else
		{
			r0dotr0 = pdot(o, lev, r0, z0);
			delta = r1dotz1 / r0dotr0;
			for(i = 0; i < npno; i++)
				s2[i] = z1[i] + delta * s1[i];
		}
		orc_grad(o, lev, s2, Ah);
		valid = orc_solve_del2_u(o, u1, Ah, imp * v_res, lev, &mgc);
		mgtot += mgc;
		orc_strip(o, lev, u1);
		orc_div(o, lev, u1, Ah);
		s2dotAhat = pdot(o, lev, s2, Ah);
		alpha = valid ? r1dotz1 / s2dotAhat : 0.0;
		for(i = 0; i < npno; i++)
		{
			r2[i] = r1[i] - alpha * Ah[i];
			P[i] += alpha * s2[i];
		}
		for(i = 0; i < neq; i++)
			V[i] -= alpha * u1[i];
		orc_div(o, lev, V, Ah);
		vdotv = vdot(o, lev, V, V);
		pdotp = pdot(o, lev, P, P);
		incomp = sqrt((gneq / gnpno) * (1.0e-32 + pdot(o, lev, Ah, Ah) / (1.0e-32 + vdotv)));
		dpressure = (float)(alpha * sqrt(pdot(o, lev, s2, s2) / (1.0e-32 + pdotp)));
		dvelocity = (float)(alpha * sqrt(vdot(o, lev, u1, u1) / (1.0e-32 + vdotv)));
		vdotv = sqrt(vdotv / gneq);
		pdotp = sqrt(pdotp / gnpno);
		count++;
		sh = s1;
		s1 = s2;
		s2 = sh;
		sh = r0;
		r0 = r1;
		r1 = r2;
		r2 = sh;
		sh = z0;
		z0 = z1;
		z1 = sh;
	}
	if(stats)
	{
		stats[0] = count;
		stats[1] = vdotv;
		stats[2] = pdotp;
		stats[3] = incomp;
		stats[4] = mgtot;
	}
	free(r0);
	free(r1);
	free(r2);
	free(z0);
	free(z1);
	free(s1);
	free(s2);
	free(Ah);
	free(u1);
	return count;
}

/* ------------------------------------------------------------------------ */
/* buoyancy, forces, energy                                                  */

/* return_horiz_ave with get_global_1d_shape_fn (cartesian, bottom face),
 * Global_operations.c:133-251, Size_does_matter.c:447-531 */
static void horiz_ave(orc *o, const float *X, float *H)
{
	const int lev = o->levmax, noz = o->noz[lev], nox = o->nox[lev];
	const int elz = o->elz[lev], elx = o->elx[lev], ely = o->ely[lev];
	double *temp = calloc(2 * noz + 2, sizeof(double));
	int i, j, k, d, nint;
	for(i = 1; i <= elz; i++)
	{
		temp[i] = temp[i + noz] = 0.0;
		temp[i + 1] = temp[i + 1 + noz] = 0.0;
		for(j = 1; j <= elx; j++)
			for(k = 1; k <= ely; k++)
			{
				/* surface Jacobian of the bottom face of element (i,j,k) */
				double dG[5], xx[3][5];
				int lnode[5], ii;
				const int b = (i - 1) + noz * ((j - 1) + nox * (k - 1));
				const int bn[5] = { 0, b, b + noz, b + noz + noz * nox, b + noz * nox };
				for(ii = 1; ii <= 4; ii++)
				{
					xx[1][ii] = o->X[lev][0][bn[ii]];
					xx[2][ii] = o->X[lev][1][bn[ii]];
				}
				for(nint = 1; nint <= 4; nint++)
				{
					double dxda[3][3] = { {0} };
					int dd, ee;
					for(ii = 1; ii <= 4; ii++)
						for(dd = 1; dd <= 2; dd++)
							for(ee = 1; ee <= 2; ee++)
								dxda[dd][ee] += xx[ee][ii] * o->Mxv[(dd - 1) * 16 + 4 * (ii - 1) + nint - 1];
					dG[nint] = dxda[1][1] * dxda[2][2] - dxda[1][2] * dxda[2][1];
				}
				lnode[1] = bn[1];
				lnode[2] = bn[2];
				lnode[3] = bn[3];
				lnode[4] = bn[4];
				for(d = 1; d <= 4; d++)
					for(nint = 1; nint <= 4; nint++)
					{
						temp[i] += X[lnode[d]] * o->Mv[4 * (d - 1) + nint - 1] * dG[nint];
						temp[i + noz] += o->Mv[4 * (d - 1) + nint - 1] * dG[nint];
					}
				if(i == elz)
				{
					for(d = 1; d <= 4; d++)
						lnode[d] = bn[d] + 1;
					for(d = 1; d <= 4; d++)
						for(nint = 1; nint <= 4; nint++)
						{
							temp[i + 1] += X[lnode[d]] * o->Mv[4 * (d - 1) + nint - 1] * dG[nint];
							temp[i + 1 + noz] += o->Mv[4 * (d - 1) + nint - 1] * dG[nint];
						}
				}
			}
	}
	for(i = 1; i <= noz; i++)
		if(temp[i + noz] != 0.0)
			H[i - 1] = (float)(temp[i] / temp[i + noz]);
	free(temp);
}

/* thermal_buoyancy, Pan_problem_misc_functions.c:83-153 (thermal only) */
void orc_thermal_buoyancy(orc *o, const float *T, float *buoy)
{
	const int lev = o->levmax, nno = o->nno[lev], noz = o->noz[lev];
	float *H = calloc(noz, sizeof(float));
	int i;
	for(i = 0; i < nno; i++)
		buoy[i] = o->p.rayleigh * T[i] * o->expansivity[i % noz];
	horiz_ave(o, buoy, H);
	for(i = 0; i < nno; i++)
		buoy[i] -= H[i % noz];
	free(H);
}

/* assemble_forces with get_elt_f (CART3D, no imposed velocities),
 * Element_calculations.c:74-125, 989-1121 */
void orc_assemble_forces(orc *o, const float *buoy, double *F)
{
	const int lev = o->levmax;
	int e, a, j, k;
	for(a = 0; a < o->neq[lev]; a++)
		F[a] = 0.0;
	for(e = 0; e < o->nel[lev]; e++)
	{
		int ez, ex, ey, nodes[9];
		double force[9], fg[9];
		const float *gda = o->GDA[lev] + (size_t)e * GDA_SZ;
		edecomp(o, lev, e, &ez, &ex, &ey);
		for(a = 1; a <= 8; a++)
		{
			nodes[a] = enode(o, lev, ez, ex, ey, a);
			force[a] = buoy[nodes[a]];
		}
		for(j = 1; j <= 8; j++)
		{
			fg[j] = 0.0;
			for(k = 1; k <= 8; k++)
				fg[j] += force[k] * o->Nv[GNV(k, j)];
		}
		for(a = 1; a <= 8; a++)
		{
			double f = 0.0;
			for(j = 1; j <= 8; j++)
				f += fg[j] * o->Nv[GNV(a, j)] * gda[j] * 1.0f;
			F[3 * nodes[a] + 2] += f;
		}
	}
	orc_strip(o, lev, F);
}

static float tmax(const orc *o, const float *T)
{
	float t = -10.0f;
	int i;
	for(i = 0; i < o->nno[o->levmax]; i++)
		t = fmaxf_(T[i], t);
	return t;
}

/* std_timestep, Advection_diffusion.c:737-810 */
static float std_timestep(orc *o, const float *V)
{
	const int lev = o->levmax, nno = o->nno[lev];
	const float *V1 = V, *V2 = V + nno, *V3 = V + 2 * nno;
	float adv = 1.0e8f;
	int e, n, d;
	if(o->p.fixed_timestep != 0.0f)
		return o->p.fixed_timestep;
	if(!o->dt_been_here)
	{
		float dts = 1.0e8f;
		for(e = 0; e < o->nel[lev]; e++)
			for(d = 1; d <= 3; d++)
			{
				const float s = o->ECO[lev][(size_t)e * ECO_SZ + 4 + d];
				const float ts = s * s;
				if(dts > ts)
					dts = ts;
			}
		o->diff_timestep = (float)(0.5 * dts);
		o->dt_been_here = 1;
	}
	for(e = 0; e < o->nel[lev]; e++)
	{
		int ez, ex, ey;
		float uc1 = 0, uc2 = 0, uc3 = 0, uc, step;
		const float *eco = o->ECO[lev] + (size_t)e * ECO_SZ;
		edecomp(o, lev, e, &ez, &ex, &ey);
		for(n = 1; n <= 8; n++)
		{
			const int nd = enode(o, lev, ez, ex, ey, n);
			uc1 = (float)(uc1 + o->Np[n - 1] * V1[nd]);
			uc2 = (float)(uc2 + o->Np[n - 1] * V2[nd]);
			uc3 = (float)(uc3 + o->Np[n - 1] * V3[nd]);
		}
		uc = (float)(fabs(uc1) / eco[5] + fabs(uc2) / eco[6] + fabs(uc3) / eco[7]);
		step = (float)(0.5 / uc);
		adv = fminf_(adv, step);
	}
	adv = (float)(1.0e-32 + fminf_(o->p.finetunedt * adv, o->diff_timestep));
	return adv;
}

/* pg_shape_fn + element_residual (CART3D, diffusion on, no FBZ flux),
 * Advection_diffusion.c:448-555, 564-727; pg_solver :398-440 */
static void pg_solver(orc *o, const float *T, const float *Tdot, float *DTdot, const float *V)
{
	const int lev = o->levmax, nno = o->nno[lev];
	const float *Vd[3] = { V, V + nno, V + 2 * nno };
	int el, a, i, j;
	for(i = 0; i < nno; i++)
		DTdot[i] = 0.0f;
	for(el = 0; el < o->nel[lev]; el++)
	{
		int ez, ex, ey, nodes[9];
		const int e = el % o->elz[lev];
		const float diff = (float)((o->diffusivity[e] + o->diffusivity[e + 1]) * 0.5);
		const float *gnx = o->GNX[lev] + (size_t)el * GNX_SZ;
		const float *gda = o->GDA[lev] + (size_t)el * GDA_SZ;
		const float *eco = o->ECO[lev] + (size_t)el * ECO_SZ;
		double PG[64], uc1 = 0, uc2 = 0, uc3 = 0, uxse, ueta, ufai, xse, eta, fai, unorm, adiff;
		const double twodiff = 2.0 * diff;
		double dT[9], tx1[9], tx2[9], tx3[9], v1[9], v2[9], v3[9], Eres[9], Q;
		edecomp(o, lev, el, &ez, &ex, &ey);
		for(a = 1; a <= 8; a++)
			nodes[a] = enode(o, lev, ez, ex, ey, a);
		/* pg_shape_fn */
		for(i = 1; i <= 8; i++)
		{
			uc1 += o->Np[i - 1] * Vd[0][nodes[i]];
			uc2 += o->Np[i - 1] * Vd[1][nodes[i]];
			uc3 += o->Np[i - 1] * Vd[2][nodes[i]];
		}
		uxse = fabs(uc1 * (double)eco[5]);
		ueta = fabs(uc2 * (double)eco[6]);
		ufai = fabs(uc3 * (double)eco[7]);
		xse = (uxse > twodiff) ? (1.0 - twodiff / uxse) : 0.0;
		eta = (ueta > twodiff) ? (1.0 - twodiff / ueta) : 0.0;
		fai = (ufai > twodiff) ? (1.0 - twodiff / ufai) : 0.0;
		unorm = uc1 * uc1 + uc2 * uc2 + uc3 * uc3;
		adiff = (unorm > 0.000001) ? ((uxse * xse + ueta * eta + ufai * fai) / (2.0 * unorm)) : 0.0;
		for(i = 1; i <= 8; i++)
		{
			double u1 = 0, u2 = 0, u3 = 0;
			for(j = 1; j <= 8; j++)
			{
				u1 += Vd[0][nodes[j]] * o->Nv[GNV(j, i)];
				u2 += Vd[1][nodes[j]] * o->Nv[GNV(j, i)];
				u3 += Vd[2][nodes[j]] * o->Nv[GNV(j, i)];
			}
			for(j = 1; j <= 8; j++)
			{
				const double prod1 = (u1 * gnx[GNVX(0, j, i)] + u2 * gnx[GNVX(1, j, i)] + u3 * gnx[GNVX(2, j, i)]);
				PG[GNV(j, i)] = o->Nv[GNV(j, i)] + adiff * prod1;
			}
		}
		/* element_residual */
		for(i = 1; i <= 8; i++)
			dT[i] = v1[i] = tx1[i] = v2[i] = tx2[i] = v3[i] = tx3[i] = 0.0;
		for(j = 1; j <= 8; j++)
		{
			const int node = nodes[j];
			const double Tn = T[node];
			const double DT = (o->NODE[lev][node] & (ORC_TBX | ORC_TBY | ORC_TBZ)) ? 0.0 : Tdot[node];
			for(i = 1; i <= 8; i++)
			{
				const double sfn = o->Nv[GNV(j, i)];
				dT[i] += DT * o->Nv[GNV(j, i)];
				tx1[i] += gnx[GNVX(0, j, i)] * Tn;
				tx2[i] += gnx[GNVX(1, j, i)] * Tn;
				tx3[i] += gnx[GNVX(2, j, i)] * Tn;
				v1[i] += Vd[0][node] * sfn;
				v2[i] += Vd[1][node] * sfn;
				v3[i] += Vd[2][node] * sfn;
			}
		}
		/* heating_adi = heating_visc = 0, heating_latent = 1 (Instructions.c:380) */
		Q = o->p.Q0;
		Q = (Q - 0.0f + 0.0f) * 1.0f;
		for(j = 1; j <= 8; j++)
		{
			Eres[j] = 0.0;
			for(i = 1; i <= 8; i++)
			{
				const float dl = diff * 1.0f * gda[i];
				Eres[j] -= PG[GNV(j, i)] * gda[i] * (dT[i] - Q + v1[i] * tx1[i] + v2[i] * tx2[i] + v3[i] * tx3[i]) + dl * (gnx[GNVX(0, j, i)] * tx1[i] + gnx[GNVX(1, j, i)] * tx2[i] + gnx[GNVX(2, j, i)] * tx3[i]);
			}
		}
		for(a = 1; a <= 8; a++)
			DTdot[nodes[a]] = (float)(DTdot[nodes[a]] + Eres[a]);
	}
	for(i = 0; i < nno; i++)
		DTdot[i] *= o->MASS[lev][i];
}

/* temperatures_conform_bcs, Boundary_conditions.c:949-990 (TBZ only: the
 * side walls carry zero-flux flags) */
static void temperatures_conform(orc *o, float *T)
{
	const int lev = o->levmax;
	int n;
	for(n = 0; n < o->nno[lev]; n++)
		if(o->NODE[lev][n] & ORC_TBZ)
			T[n] = o->TB[2][n];
}

/* PG_timestep, Advection_diffusion.c:251-345 with predictor :352-366 and
 * corrector :369-387.  stats[0] = T_interior, [1] = sub-iterations used */
float orc_pg_timestep(orc *o, float *T, float *Tdot, const float *V, float *stats)
{
	const int lev = o->levmax, nno = o->nno[lev];
	float *DTdot = malloc(sizeof(float) * nno), *T1 = malloc(sizeof(float) * nno), *Tdot1 = malloc(sizeof(float) * nno);
	float timestep = std_timestep(o, V), Tint1, Tint = 0, dt_reduced = 1.0f;
	int last_sub = 1, iredo, i, pass;
	for(i = 0; i < nno; i++)
	{
		T1[i] = T[i];
		Tdot1[i] = Tdot[i];
	}
	Tint1 = tmax(o, T);
	do
	{
		float mult;
		timestep *= dt_reduced;
		iredo = 0;
		mult = (float)((1.0 - o->p.gamma) * timestep);
		for(i = 0; i < nno; i++)
		{
			if(!(o->NODE[lev][i] & (ORC_TBX | ORC_TBZ | ORC_TBY)))
				T[i] += mult * Tdot[i];
			Tdot[i] = 0.0f;
		}
		for(pass = 0; pass < o->p.temp_iterations; pass++)
		{
			pg_solver(o, T, Tdot, DTdot, V);
			mult = o->p.gamma * timestep;
			for(i = 0; i < nno; i++)
			{
				if(!(o->NODE[lev][i] & (ORC_TBX | ORC_TBZ | ORC_TBY)))
					T[i] += mult * DTdot[i];
				Tdot[i] += DTdot[i];
			}
		}
		Tint = tmax(o, T);
		if(Tint / Tint1 > 1.01f)
		{
			for(i = 0; i < nno; i++)
			{
				T[i] = T1[i];
				Tdot[i] = Tdot1[i];
			}
			iredo = 1;
			dt_reduced *= 0.5f;
			last_sub++;
		}
	} while(iredo == 1 && last_sub <= 5);
	o->elapsed_time += timestep;
	temperatures_conform(o, T);
	if(stats)
	{
		stats[0] = Tint;
		stats[1] = (float)last_sub;
		stats[2] = o->elapsed_time;
	}
	free(DTdot);
	free(T1);
	free(Tdot1);
	return timestep;
}

/* heat_flux, Process_buoyancy.c:63-203: nu[0] = Nu_top, nu[1] = Nu_bot */
/* volume-RMS velocity, sqrt(int v.v dV / int dV), 2x2x2 Gauss quadrature
 * with the shape functions and dA of the finest level (no routine in the
 * reference; the definition both sides use) */
double orc_vrms(orc *o, const float *V)
{
	const int lev = o->levmax, nno = o->nno[lev];
	double vv = 0.0, vol = 0.0;
	int e, i, j, d;
	for(e = 0; e < o->nel[lev]; e++)
	{
		int ez, ex, ey, nodes[9];
		const float *gda = o->GDA[lev] + (size_t)e * GDA_SZ;
		edecomp(o, lev, e, &ez, &ex, &ey);
		for(j = 1; j <= 8; j++)
			nodes[j] = enode(o, lev, ez, ex, ey, j);
		for(i = 1; i <= 8; i++)
		{
			double s2 = 0.0;
			for(d = 0; d < 3; d++)
			{
				double u = 0.0;
				for(j = 1; j <= 8; j++)
					u += (double)V[d * nno + nodes[j]] * o->Nv[GNV(j, i)];
				s2 += u * u;
			}
			vv += s2 * gda[i];
			vol += gda[i];
		}
	}
	return sqrt(vv / vol);
}

void orc_heat_flux(orc *o, const float *T, const float *V, float *nu)
{
	const int lev = o->levmax, nno = o->nno[lev], noz = o->noz[lev], nox = o->nox[lev], noy = o->noy[lev];
	float *hf = calloc(nno, sizeof(float));
	double hft = 0, hfb = 0, areat = 0, areab = 0;
	int e, i, j, a;
	for(e = 0; e < o->nel[lev]; e++)
	{
		int ez, ex, ey, nodes[9];
		const int ee = e % o->elz[lev];
		const double diff = (o->diffusivity[ee] + o->diffusivity[ee + 1]) * 0.5;
		const float *gnx = o->GNX[lev] + (size_t)e * GNX_SZ;
		const float *gda = o->GDA[lev] + (size_t)e * GDA_SZ;
		double VZ[9], uT = 0.0;
		edecomp(o, lev, e, &ez, &ex, &ey);
		for(j = 1; j <= 8; j++)
		{
			nodes[j] = enode(o, lev, ez, ex, ey, j);
			VZ[j] = V[2 * nno + nodes[j]];
		}
		for(i = 1; i <= 8; i++)
		{
			double u = 0, Tg = 0, dTdz = 0;
			for(j = 1; j <= 8; j++)
			{
				u += VZ[j] * o->Nv[GNV(j, i)];
				Tg += T[nodes[j]] * o->Nv[GNV(j, i)];
				dTdz = dTdz + T[nodes[j]] * gnx[GNVX(2, j, i)];
			}
			uT = uT + (u * Tg - diff * dTdz) * gda[i];
		}
		uT /= o->ECO[lev][(size_t)e * ECO_SZ + 12];
		for(j = 1; j <= 8; j++)
			hf[nodes[j]] = (float)(hf[nodes[j]] + o->TWW[lev][(size_t)e * TWW_SZ + j] * uT);
	}
	for(i = 0; i < nno; i++)
		hf[i] = hf[i] * o->MASS[lev][i];
	/* surface extrapolation and surface-element sums (:158-193) */
	{
		float *sh = calloc(nox * noy, sizeof(float)), *bh = calloc(nox * noy, sizeof(float));
		for(i = 0; i < nox * noy; i++)
		{
			const int top = i * noz + noz - 1, bot = i * noz;
			sh[i] = 2 * hf[top] - hf[top - 1];
			bh[i] = 2 * hf[bot] - hf[bot + 1];
		}
		for(j = 0; j < o->ely[lev]; j++)
			for(i = 0; i < o->elx[lev]; i++)
			{
				/* sien: surface nodes of the top face, in surface numbering */
				const int s1 = i + nox * j, s2 = s1 + 1, s3 = s1 + 1 + nox, s4 = s1 + nox;
				/* the surface numbering of the reference is x-fastest over
				 * (x, y) because surf_node[i] = i*noz; node/noz - 1 = x + nox*y */
				const double tb = (double)bh[s1] + bh[s2] + bh[s3] + bh[s4];
				const double tt = (double)sh[s1] + sh[s2] + sh[s3] + sh[s4];
				const int eb = o->elz[lev] * (i + o->elx[lev] * j), et = eb + o->elz[lev] - 1;
				(void)a;
				hfb += tb * o->ECO[lev][(size_t)eb * ECO_SZ + 12];
				areab += o->ECO[lev][(size_t)eb * ECO_SZ + 12];
				hft += tt * o->ECO[lev][(size_t)et * ECO_SZ + 12];
				areat += o->ECO[lev][(size_t)et * ECO_SZ + 12];
			}
		{
			const float inp[4] = { (float)hfb, (float)areab, (float)hft, (float)areat };
			nu[1] = inp[0] / (inp[1] * 4);
			nu[0] = inp[2] / (inp[3] * 4);
		}
		free(sh);
		free(bh);
	}
	free(hf);
}
