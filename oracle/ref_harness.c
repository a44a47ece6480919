/*
 * ref_harness -- golden-vector generator linked against the UNMODIFIED
 * reference objects (everything but Citcom.o).  Test infrastructure only.
 *
 *   ref_harness <deck> <outdir> setup
 *       state after the initialisation of Instructions.c:55-157 and the
 *       operator build of general_stokes_solver (Drive_solvers.c:105-126):
 *       per-level geometry, flags, viscosity, stiffness (Eqn_k1..3,
 *       Node_map), BI, BPI, elt_del, TWW, MASS, and the finest-level fields.
 *   ref_harness <deck> <outdir> ops <seed>
 *       seeded random inputs -> n_assemble_del2_u, gauss_seidel,
 *       project_vector, interp_vector, assemble_div_u, assemble_grad_p,
 *       strip_bcs_from_residual on every level (inputs and outputs dumped).
 *   ref_harness <deck> <outdir> run <nsteps>
 *       the time loop of Citcom.c:97-161 for nsteps steps; per-step scalars
 *       to steps.txt and final T, Tdot, U, P, V, Nu and the layer profiles
 *       of averages (Have.T, Have.vrms, Have.Vi) to outdir.
 *   ref_harness <deck> <outdir> time
 *       a bounded timing sample of the reference on the deck (CPU baseline):
 *       the operator build of one step, one multi_grid call and one
 *       PG_timestep, printed as "build <s> mg <s> pg <s>".
 *
 * Every array is written raw (native endianness) to <outdir>/<name> (under
 * MPI with several ranks, each rank's local arrays to <outdir>/<name>.r<rank>);
 * the 1-based reference arrays are written from index 1 unless noted.
 */
#include <math.h>
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "element_definitions.h"
#include "global_defs.h"

extern int Emergency_stop;

static const char *g_out;
static int g_rank = 0, g_nranks = 1;

/* under MPI with several ranks every rank writes its own local arrays, the
 * file name suffixed with .r<rank> */
static void dump(const char *name, const void *p, size_t bytes)
{
	char path[1024];
	FILE *fp;
	if(g_nranks > 1)
		snprintf(path, sizeof path, "%s/%s.r%d", g_out, name, g_rank);
	else
		snprintf(path, sizeof path, "%s/%s", g_out, name);
	fp = fopen(path, "wb");
	if(!fp)
	{
		perror(path);
		exit(2);
	}
	if(bytes && fwrite(p, 1, bytes, fp) != bytes)
	{
		perror(path);
		exit(2);
	}
	fclose(fp);
}

static void dumpl(const char *base, int lev, const void *p, size_t bytes)
{
	char name[256];
	snprintf(name, sizeof name, "%s_lev%d", base, lev);
	dump(name, p, bytes);
}

/* deterministic uniform(-1,1) generator, independent of libc */
static unsigned long long g_rng = 88172645463325252ULL;
static double urand(void)
{
	g_rng ^= g_rng << 13;
	g_rng ^= g_rng >> 7;
	g_rng ^= g_rng << 17;
	return ((double)(g_rng >> 11) / 9007199254740992.0) * 2.0 - 1.0;
}

static void dump_setup(struct All_variables *E)
{
	int lev, d;
	int dims[3 * MAX_LEVELS + 2];
	dims[0] = E->mesh.levmin;
	dims[1] = E->mesh.levmax;
	for(lev = 0; lev < MAX_LEVELS; lev++)
	{
		dims[2 + 3 * lev] = E->lmesh.NOX[lev];
		dims[3 + 3 * lev] = E->lmesh.NOY[lev];
		dims[4 + 3 * lev] = E->lmesh.NOZ[lev];
	}
	dump("dims.i32", dims, sizeof dims);

	for(lev = E->mesh.levmin; lev <= E->mesh.levmax; lev++)
	{
		const int nno = E->lmesh.NNO[lev], nel = E->lmesh.NEL[lev], neq = E->lmesh.NEQ[lev];
		const int max_eqn = max_eqn_interaction[3];
		char nm[64];
		for(d = 1; d <= 3; d++)
		{
			snprintf(nm, sizeof nm, "X%d.f32", d);
			dumpl(nm, lev, E->XX[lev][d] + 1, nno * sizeof(float));
		}
		dumpl("NODE.u32", lev, E->NODE[lev] + 1, nno * sizeof(unsigned int));
		dumpl("GNX.f32", lev, E->GNX[lev] + 1, nel * sizeof(struct Shape_function_dx));
		dumpl("GDA.f32", lev, E->GDA[lev] + 1, nel * sizeof(struct Shape_function_dA));
		dumpl("EVI.f32", lev, E->EVI[lev] + 1, nel * 8 * sizeof(float));
		dumpl("MASS.f32", lev, E->MASS[lev] + 1, nno * sizeof(float));
		dumpl("TWW.f32", lev, E->TWW[lev] + 1, nel * sizeof(struct FNODE));
		dumpl("ECO.f32", lev, E->ECO[lev] + 1, nel * sizeof(struct COORD));
		dumpl("BI.f64", lev, E->BI[lev], neq * sizeof(double));
		dumpl("BPI.f64", lev, E->BPI[lev] + 1, E->lmesh.NPNO[lev] * sizeof(double));
		dumpl("EG.f32", lev, E->elt_del[lev] + 1, nel * sizeof(struct EG));
		dumpl("K1.f32", lev, E->Eqn_k1[lev], (size_t)nno * max_eqn * sizeof(float));
		dumpl("K2.f32", lev, E->Eqn_k2[lev], (size_t)nno * max_eqn * sizeof(float));
		dumpl("K3.f32", lev, E->Eqn_k3[lev], (size_t)nno * max_eqn * sizeof(float));
		dumpl("MAP.i32", lev, E->Node_map[lev], (size_t)nno * max_eqn * sizeof(int));
	}
	{
		const int nno = E->lmesh.nno;
		dump("T.f32", E->T + 1, nno * sizeof(float));
		dump("buoyancy.f32", E->buoyancy + 1, nno * sizeof(float));
		dump("F.f64", E->F, E->lmesh.neq * sizeof(double));
		dump("Mass.f32", E->Mass + 1, nno * sizeof(float));
		dump("diffusivity.f32", E->diffusivity + 1, E->lmesh.noz * sizeof(float));
		dump("expansivity.f32", E->expansivity + 1, E->lmesh.noz * sizeof(float));
		dump("node.u32", E->node + 1, nno * sizeof(unsigned int));
		for(d = 1; d <= 3; d++)
		{
			char nm[32];
			snprintf(nm, sizeof nm, "TB%d.f32", d);
			dump(nm, E->TB[d] + 1, nno * sizeof(float));
			snprintf(nm, sizeof nm, "VB%d.f32", d);
			dump(nm, E->VB[d] + 1, nno * sizeof(float));
		}
	}
}

/* the operator build that general_stokes_solver performs before the solve
 * (Drive_solvers.c:105-126) */
static void build_operators(struct All_variables *E)
{
	velocities_conform_bcs(E, E->U);
	assemble_forces(E, 0);
	if(E->viscosity.update_allowed)
		get_system_viscosity(E, 1, E->EVI[E->mesh.levmax], E->VI[E->mesh.levmax]);
	construct_stiffness_B_matrix(E);
}

static void fill_rand(double *v, int n)
{
	int i;
	for(i = 0; i < n; i++)
		v[i] = urand();
}

static void dump_ops(struct All_variables *E, unsigned long long seed)
{
	int lev;
	g_rng ^= seed * 0x9E3779B97F4A7C15ULL;
	for(lev = E->mesh.levmax; lev >= E->mesh.levmin; lev--)
	{
		const int neq = E->lmesh.NEQ[lev], npno = E->lmesh.NPNO[lev];
		double *u = calloc(neq + 2, sizeof(double));
		double *Au = calloc(neq + 2, sizeof(double));
		double *F = calloc(neq + 2, sizeof(double));
		double *d0 = calloc(neq + 2, sizeof(double));
		double *Ad = calloc(neq + 2, sizeof(double));
		double *p = calloc(npno + 2, sizeof(double));
		double *div = calloc(npno + 2, sizeof(double));
		double *grad = calloc(neq + 2, sizeof(double));
		int cycles;

		/* matvec (Element_calculations.c:552-621), strip_bcs = 1 */
		fill_rand(u, neq);
		dumpl("mv_in.f64", lev, u, neq * sizeof(double));
		n_assemble_del2_u(E, u, Au, lev, 1);
		dumpl("mv_out.f64", lev, Au, neq * sizeof(double));

		/* gauss_seidel from zero (guess = 0), 3 sweeps */
		fill_rand(F, neq);
		strip_bcs_from_residual(E, F, lev);
		dumpl("gs0_F.f64", lev, F, neq * sizeof(double));
		cycles = 3;
		gauss_seidel(E, d0, F, Ad, 0.01, &cycles, lev, 0);
		dumpl("gs0_d.f64", lev, d0, neq * sizeof(double));
		dumpl("gs0_Ad.f64", lev, Ad, neq * sizeof(double));

		/* gauss_seidel from a given guess (guess = 1), 2 sweeps */
		fill_rand(d0, neq);
		strip_bcs_from_residual(E, d0, lev);
		dumpl("gs1_d_in.f64", lev, d0, neq * sizeof(double));
		cycles = 2;
		gauss_seidel(E, d0, F, Ad, 0.01, &cycles, lev, 1);
		dumpl("gs1_d.f64", lev, d0, neq * sizeof(double));
		dumpl("gs1_Ad.f64", lev, Ad, neq * sizeof(double));

		/* div and grad (Element_calculations.c:691-769) */
		assemble_div_u(E, u, div, lev);
		dumpl("div_out.f64", lev, div + 1, npno * sizeof(double));
		fill_rand(p + 1, npno);
		p[1 + npno / 3] = 0.0;	/* exercises the P[e]==0 skip */
		dumpl("grad_in.f64", lev, p + 1, npno * sizeof(double));
		assemble_grad_p(E, p, grad, lev);
		dumpl("grad_out.f64", lev, grad, neq * sizeof(double));

		if(lev > E->mesh.levmin)
		{
			const int neqc = E->lmesh.NEQ[lev - 1];
			double *c = calloc(neqc + 2, sizeof(double));
			double *f = calloc(neq + 2, sizeof(double));
			/* restriction (Solver_multigrid.c:72-159) of the random u */
			project_vector(E, lev, u, c, 1);
			dumpl("proj_out.f64", lev, c, neqc * sizeof(double));
			/* prolongation (Solver_multigrid.c:173-298) of a random coarse vector */
			fill_rand(c, neqc);
			dumpl("interp_in.f64", lev, c, neqc * sizeof(double));
			interp_vector(E, lev - 1, c, f);
			dumpl("interp_out.f64", lev, f, neq * sizeof(double));
			free(c);
			free(f);
		}
		free(u);
		free(Au);
		free(F);
		free(d0);
		free(Ad);
		free(p);
		free(div);
		free(grad);
	}
}

static void dump_fields(struct All_variables *E, const char *tag)
{
	char nm[64];
	const int nno = E->lmesh.nno;
	int d;
	snprintf(nm, sizeof nm, "%s_T.f32", tag);
	dump(nm, E->T + 1, nno * sizeof(float));
	snprintf(nm, sizeof nm, "%s_Tdot.f32", tag);
	dump(nm, E->Tdot + 1, nno * sizeof(float));
	snprintf(nm, sizeof nm, "%s_U.f64", tag);
	dump(nm, E->U, E->lmesh.neq * sizeof(double));
	snprintf(nm, sizeof nm, "%s_P.f64", tag);
	dump(nm, E->P + 1, E->lmesh.npno * sizeof(double));
	for(d = 1; d <= 3; d++)
	{
		snprintf(nm, sizeof nm, "%s_V%d.f32", tag, d);
		dump(nm, E->V[d] + 1, nno * sizeof(float));
	}
	snprintf(nm, sizeof nm, "%s_buoyancy.f32", tag);
	dump(nm, E->buoyancy + 1, nno * sizeof(float));
}

/* time loop of Citcom.c:97-161 with per-step scalars recorded */
static void run_steps(struct All_variables *E, int nsteps)
{
	char path[1024];
	FILE *fp;
	snprintf(path, sizeof path, "%s/steps.txt", g_out);
	/* the scalars are global (horizontal sums over ranks): rank 0 writes them */
	fp = fopen(g_rank == 0 ? path : "/dev/null", "w");

	general_stokes_solver(E);
	process_temp_field(E, E->monitor.solution_cycles);
	heat_flux(E);
	fprintf(fp, "%d %.9e %.9e %.9e %.9e %.9e\n", E->monitor.solution_cycles, E->monitor.elapsed_time, E->advection.timestep, E->slice.Nut, E->slice.Nub, E->monitor.T_interior);
	dump_fields(E, "step0");
	process_new_velocity(E, E->monitor.solution_cycles);

	while(E->monitor.solution_cycles < nsteps && Emergency_stop == 0)
	{
		process_heating(E);
		E->monitor.solution_cycles++;
		if(E->monitor.solution_cycles > E->control.print_convergence)
			E->control.print_convergence = 1;
		(E->next_buoyancy_field) (E);
		if(E->monitor.solution_cycles == 1)
			dump_fields(E, "pg1");
		process_temp_field(E, E->monitor.solution_cycles);
		general_stokes_solver(E);
		process_new_velocity(E, E->monitor.solution_cycles);
		heat_flux(E);
		fprintf(fp, "%d %.9e %.9e %.9e %.9e %.9e\n", E->monitor.solution_cycles, E->monitor.elapsed_time, E->advection.timestep, E->slice.Nut, E->slice.Nub, E->monitor.T_interior);
		fflush(fp);
	}
	fclose(fp);
	dump_fields(E, "final");
	/* the layer profiles of the .ave output after the last step: Have.T from
	 * heat_flux (Process_buoyancy.c:93), Have.vrms and Have.Vi from averages
	 * (Process_velocity.c:179-224) */
	averages(E);
	dump("final_aveT.f32", E->Have.T + 1, E->lmesh.noz * sizeof(float));
	dump("final_avevrms.f32", E->Have.vrms + 1, E->lmesh.noz * sizeof(float));
	dump("final_aveVi.f32", E->Have.Vi + 1, E->lmesh.noz * sizeof(float));
}

/* the reference's own routines, timed with its own clock (CPU_time0,
 * MPI_Wtime): build_operators as general_stokes_solver does it
 * (Drive_solvers.c:105-126), multi_grid (General_matrix_functions.c:525)
 * on the assembled forces, PG_timestep (Advection_diffusion.c:251) */
static void time_sample(struct All_variables *E)
{
	const int neq = E->lmesh.neq;
	double *d = calloc(neq + 2, sizeof(double)), *F = calloc(neq + 2, sizeof(double));
	double *Au = calloc(neq + 2, sizeof(double));
	double t0, t_build, t_mg, t_pg;
	t0 = CPU_time0();
	build_operators(E);
	t_build = CPU_time0() - t0;
	memcpy(F, E->F, neq * sizeof(double));
	t0 = CPU_time0();
	multi_grid(E, d, F, Au, E->control.accuracy, E->mesh.levmax);
	t_mg = CPU_time0() - t0;
	t0 = CPU_time0();
	PG_timestep(E);
	t_pg = CPU_time0() - t0;
	printf("build %.6f mg %.6f pg %.6f\n", t_build, t_mg, t_pg);
	free(d);
	free(F);
	free(Au);
}

int main(int argc, char **argv)
{
	struct All_variables E;
	const char *mode;

	if(argc < 4)
	{
		fprintf(stderr, "usage: %s deck outdir setup|ops <seed>|run <nsteps>|time\n", argv[0]);
		return 2;
	}
	g_out = argv[2];
	mode = argv[3];
	mkdir(g_out, 0755);

	srand48((long int)-1);	/* as Citcom.c:58 */
	memset(&E, 0, sizeof E);
	E.parallel.me = 0;
	E.parallel.nproc = 1;
	MPI_Init(&argc, &argv);
	MPI_Comm_rank(MPI_COMM_WORLD, &(E.parallel.me));
	MPI_Comm_size(MPI_COMM_WORLD, &(E.parallel.nproc));
	g_rank = E.parallel.me;
	g_nranks = E.parallel.nproc;
	E.monitor.solution_cycles = 0;
	E.monitor.elapsed_time = 0;
	E.advection.timestep = 0;
	E.advection.timesteps = 0;

	read_instructions(&E, 2, argv);
	E.control.keep_going = 1;

	if(strcmp(mode, "setup") == 0)
	{
		build_operators(&E);
		dump_setup(&E);
	}
	else if(strcmp(mode, "ops") == 0)
	{
		build_operators(&E);
		dump_ops(&E, argc > 4 ? strtoull(argv[4], NULL, 10) : 1ULL);
	}
	else if(strcmp(mode, "run") == 0)
	{
		run_steps(&E, argc > 4 ? atoi(argv[4]) : 1);
	}
	else if(strcmp(mode, "time") == 0)
	{
		time_sample(&E);
	}
	else
	{
		fprintf(stderr, "unknown mode %s\n", mode);
		return 2;
	}
	MPI_Finalize();
	return 0;
}
