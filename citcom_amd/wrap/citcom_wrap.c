/*
 * citcom_wrap.c -- link-time drop-in of libcitcom_hip.so into an unmodified
 * CitcomCU build.
 *
 * Linked with -Wl,--wrap=solve_constrained_flow_iterative
 *             -Wl,--wrap=PG_timestep
 *             -Wl,--wrap=construct_stiffness_B_matrix
 *             -Wl,--wrap=get_system_viscosity
 *             -Wl,--wrap=assemble_forces
 * the driver's call of solve_constrained_flow_iterative (Drive_solvers.c:130)
 * and the PG_timestep pointer installed as E->next_buoyancy_field
 * (Convection.c:93, called at Citcom.c:125) reach the functions below, which
 * hand E's arrays to the GPU through the C ABI of citcom_hip.h and write the
 * results back where the reference writes them.  The driver's host rebuild of
 * the operators (Drive_solvers.c:126), which only the replaced solver reads
 * after the first step, runs once; the device rebuilds them.  Once the
 * device holds the temperature (after the first PG_timestep), the viscosity
 * (get_system_viscosity, Drive_solvers.c:124) and the buoyancy forces
 * (assemble_forces, Drive_solvers.c:119) are formed on the device from it;
 * the viscosity is copied back for the host's readers (topography,
 * Topo_gravity.c:401; averages, Process_velocity.c:189), the forces stay on
 * the device (only the solver reads E->F).  __real_* stay linked and run
 * where the device cannot stand in (before the device holds the
 * temperature, for another array than the finest-level viscosity, or for
 * viscosity laws the device does not form).  averages itself
 * (Process_velocity.c:179-224) cannot be redirected: its caller,
 * process_new_velocity, is in the same translation unit, and --wrap only
 * redirects calls between objects; it reads the V and viscosity written
 * back here (chip_averages computes the same profiles on the device).
 *
 * CITCOM_HIP_FORM=mixed (or mixed16) in the environment selects the mixed
 * precision form (chip_problem.mixed = 1, 2); the default is the reference's
 * arithmetic.  CITCOM_HIP_WARM=extrap starts each Uzawa solve from the last
 * solutions extrapolated along the sequence of solves (chip option
 * "warm_start" = 1); the default is the reference's start.
 *
 * Compiled against the reference's own headers with the reference's -D flags
 * (struct All_variables must have the driver's layout, global_defs.h).
 * Single rank only: the library holds one subdomain (nproc = 1).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "element_definitions.h"
#include "global_defs.h"

#include "citcom_hip.h"

void parallel_process_termination(void);	/* Parallel_related.c:68-73 */
double CPU_time0(void);	/* Profiling.c */

static chip_ctx *ctx;
static float *vpack;	/* V[d*nno + n] staging for the ABI */
static int stokes_calls;
static int t_on_device;	/* the device T is E->T (after a PG_timestep) */
static int evi_on_device;	/* the device formed this step's viscosity */
static int f_on_device;	/* the device formed this step's forces */

static void fail(struct All_variables *E, const char *what)
{
	fprintf(stderr, "citcom_hip: %s failed: %s\n", what, chip_last_error());
	if(E->fp)
	{
		fprintf(E->fp, "citcom_hip: %s failed: %s\n", what, chip_last_error());
		fflush(E->fp);
	}
	parallel_process_termination();
}

#define CHECK(call) do { if((call) != 0) fail(E, #call); } while(0)

static void release(void)
{
	if(ctx)
		chip_destroy(ctx);
	ctx = NULL;
	free(vpack);
	vpack = NULL;
}

/* chip_problem from the control data read_instructions left in E
 * (Instructions.c:55-157, Viscosity_structures.c:100-320,
 * Advection_diffusion.c:70-90) and the finest-level node coordinates
 * (node_locations, Nodal_mesh.c:53-283). */
static void setup(struct All_variables *E)
{
	chip_problem p;
	const int lv = E->mesh.levmax;
	const int nox = E->lmesh.NOX[lv], noy = E->lmesh.NOY[lv], noz = E->lmesh.NOZ[lv];
	float *xc, *yc, *zc;
	int i;

	if(E->parallel.nproc != 1)
	{
		fprintf(stderr, "citcom_hip: one MPI rank per run (nproc = %d)\n", E->parallel.nproc);
		parallel_process_termination();
	}
	xc = malloc(sizeof(float) * nox);
	yc = malloc(sizeof(float) * noy);
	zc = malloc(sizeof(float) * noz);
	for(i = 0; i < nox; i++)
		xc[i] = E->X[1][1 + i * noz];
	for(i = 0; i < noy; i++)
		yc[i] = E->X[2][1 + i * noz * nox];
	for(i = 0; i < noz; i++)
		zc[i] = E->X[3][1 + i];

	memset(&p, 0, sizeof p);
	p.levels = E->mesh.levels;
	p.mgunitx = E->mesh.mgunitx;
	p.mgunity = E->mesh.mgunity;
	p.mgunitz = E->mesh.mgunitz;
	p.xc = xc;
	p.yc = yc;
	p.zc = zc;
	p.topvbc = E->mesh.topvbc;
	p.botvbc = E->mesh.botvbc;
	p.toptbc = E->mesh.toptbc;
	p.bottbc = E->mesh.bottbc;
	p.toptbcval = E->control.TBCtopval;
	p.bottbcval = E->control.TBCbotval;
	p.accuracy = E->control.accuracy;
	p.p_iterations = E->control.p_iterations;
	p.v_steps_low = E->control.v_steps_low;
	p.v_steps_high = E->control.v_steps_high;
	p.down_heavy = E->control.down_heavy;
	p.up_heavy = E->control.up_heavy;
	p.mg_cycle = E->control.mg_cycle;
	p.aug_lagr = E->control.augmented_Lagr;
	p.aug_number = E->control.augmented;
	p.precond = E->control.precondition;
	p.rheol = E->viscosity.RHEOL;
	p.tdepv = E->viscosity.TDEPV;
	p.num_mat = E->viscosity.num_mat;
	p.visc_update = E->viscosity.update_allowed;
	for(i = 0; i < 4 && i < E->viscosity.num_mat; i++)
	{
		p.visc0[i] = E->viscosity.N0[i];
		p.viscE[i] = E->viscosity.E[i];
		p.viscT[i] = E->viscosity.T[i];
		p.z_layer[i] = E->viscosity.zbase_layer[i];
	}
	p.vmin_on = E->viscosity.MIN;
	p.vmax_on = E->viscosity.MAX;
	p.visc_min = E->viscosity.min_value;
	p.visc_max = E->viscosity.max_value;
	p.rayleigh = E->control.Atemp;
	p.finetunedt = E->advection.fine_tune_dt;
	p.temp_iterations = E->advection.temp_iterations;
	p.gamma = E->advection.gamma;
	p.Q0 = E->control.Q0;
	p.fixed_timestep = E->advection.fixed_timestep;
	p.mixed = 0;
	{
		const char *form = getenv("CITCOM_HIP_FORM");
		if(form && strcmp(form, "mixed") == 0)
			p.mixed = 1;
		else if(form && strcmp(form, "mixed16") == 0)
			p.mixed = 2;
		else if(form && *form && strcmp(form, "plain") != 0)
		{
			fprintf(stderr, "citcom_hip: CITCOM_HIP_FORM=%s (plain, mixed or mixed16)\n", form);
			parallel_process_termination();
		}
	}

	CHECK(chip_create(&p, &ctx));
	{
		const char *warm = getenv("CITCOM_HIP_WARM");
		if(warm && strcmp(warm, "extrap") == 0)
			CHECK(chip_set_option(ctx, "warm_start", 1));
		else if(warm && *warm && strcmp(warm, "previous") != 0)
		{
			fprintf(stderr, "citcom_hip: CITCOM_HIP_WARM=%s (previous or extrap)\n", warm);
			parallel_process_termination();
		}
	}
	free(xc);
	free(yc);
	free(zc);
	vpack = malloc(sizeof(float) * 3 * (size_t)E->lmesh.nno);
	atexit(release);
}

void __real_construct_stiffness_B_matrix(struct All_variables *E);

/* Construct_arrays.c:834-889.  The host operators are built once: besides
 * the solver (replaced here), the reference reads only the geometry-only
 * elt_del they include (topography output, Topo_gravity.c:157).  Every later
 * rebuild happens on the device in the solve below. */
void __wrap_construct_stiffness_B_matrix(struct All_variables *E)
{
	static int built;
	if(!built)
		__real_construct_stiffness_B_matrix(E);
	built = 1;
}

/* Stokes_flow_Incomp.c:50-90 */
void __wrap_solve_constrained_flow_iterative(struct All_variables *E)
{
	const int nno = E->lmesh.nno, neq = E->lmesh.neq, npno = E->lmesh.npno, nel = E->lmesh.nel;
	chip_stokes_stats st;
	double t0 = CPU_time0();
	int d;

	if(!ctx)
		setup(E);
	/* operators: rebuilt from the viscosity the driver just computed
	 * (get_system_viscosity, Drive_solvers.c:122), as
	 * construct_stiffness_B_matrix does on the host */
	if(stokes_calls == 0 || E->viscosity.update_allowed)
	{
		if(!evi_on_device)
			CHECK(chip_upload(ctx, "EVI", -1, E->EVI[E->mesh.levmax] + 1, 8L * nel));
		CHECK(chip_construct_stiffness_B_matrix(ctx));
	}
	evi_on_device = 0;
	stokes_calls++;
	if(!f_on_device)
		CHECK(chip_upload(ctx, "F", 0, E->F, neq));
	f_on_device = 0;
	CHECK(chip_upload(ctx, "U", 0, E->U, neq));
	CHECK(chip_upload(ctx, "P", 0, E->P + 1, npno));
	CHECK(chip_solve_constrained_flow_iterative(ctx, &st));
	CHECK(chip_download(ctx, "U", 0, E->U, neq));
	CHECK(chip_download(ctx, "P", 0, E->P + 1, npno));
	CHECK(chip_download(ctx, "V", 0, vpack, 3L * nno));
	for(d = 0; d < 3; d++)
		memcpy(E->V[d + 1] + 1, vpack + (size_t)d * nno, sizeof(float) * nno);

	/* monitors and counters (Stokes_flow_Incomp.c:438-446,
	 * General_matrix_functions.c:516-517) */
	E->monitor.vdotv = st.vdotv;
	E->monitor.pdotp = st.pdotp;
	E->monitor.incompressibility = st.incompressibility;
	E->control.total_iteration_cycles += st.mg_cycles;
	E->control.total_v_solver_calls += st.inner_solves;
	if(E->control.print_convergence && E->parallel.me == 0)
	{
		fprintf(E->fp, "after (%03d) pressure loops and %g sec for step %d\n", st.iterations, CPU_time0() - t0, E->monitor.solution_cycles);
		fprintf(stderr, "after (%03d) pressure loops and %g sec for step %d\n", st.iterations, CPU_time0() - t0, E->monitor.solution_cycles);
		fflush(E->fp);
	}
}

/* Advection_diffusion.c:251-345 (with thermal_buoyancy at its end) */
void __wrap_PG_timestep(struct All_variables *E)
{
	const int nno = E->lmesh.nno;
	chip_pg_stats st;
	int d;

	if(!ctx)
		setup(E);
	E->advection.timesteps++;
	for(d = 0; d < 3; d++)
		memcpy(vpack + (size_t)d * nno, E->V[d + 1] + 1, sizeof(float) * nno);
	CHECK(chip_upload(ctx, "T", 0, E->T + 1, nno));
	CHECK(chip_upload(ctx, "Tdot", 0, E->Tdot + 1, nno));
	CHECK(chip_upload(ctx, "V", 0, vpack, 3L * nno));
	CHECK(chip_PG_timestep(ctx, &st));
	CHECK(chip_download(ctx, "T", 0, E->T + 1, nno));
	CHECK(chip_download(ctx, "Tdot", 0, E->Tdot + 1, nno));
	CHECK(chip_download(ctx, "buoyancy", 0, E->buoyancy + 1, nno));
	t_on_device = 1;

	E->advection.timestep = st.timestep;
	E->advection.dt_reduced = 1.0f;
	for(d = 1; d < st.sub_iterations; d++)
		E->advection.dt_reduced *= 0.5f;
	E->monitor.T_interior = st.T_interior;
	E->advection.total_timesteps++;
	E->monitor.elapsed_time += E->advection.timestep;
	E->advection.last_sub_iterations = 1;
	E->control.keep_going = (E->monitor.solution_cycles < E->advection.max_timesteps) ? 1 : 0;
}

void __real_get_system_viscosity(struct All_variables *E, int propogate, float *evisc, float *visc);
void __real_assemble_forces(struct All_variables *E, int penalty);

/* Viscosity_structures.c:369-443: the finest level's Gauss-point viscosity
 * from the device temperature (visc_from_T / visc_from_mat and the MIN/MAX
 * clip, as chip_get_system_viscosity), copied back into evisc for the
 * host's other readers */
void __wrap_get_system_viscosity(struct All_variables *E, int propogate, float *evisc, float *visc)
{
	if(!ctx || !t_on_device || evisc != E->EVI[E->mesh.levmax] || E->viscosity.SMOOTH || E->viscosity.SDEPV ||
	   E->viscosity.CDEPV || E->viscosity.BDEPV || E->control.KERNEL != 1)
	{
		__real_get_system_viscosity(E, propogate, evisc, visc);
		return;
	}
	CHECK(chip_get_system_viscosity(ctx));
	CHECK(chip_download(ctx, "EVI", -1, evisc + 1, 8L * E->lmesh.nel));
	evi_on_device = 1;
}

/* Element_calculations.c:74-125: the buoyancy forces from the device
 * buoyancy (the PG step left it there); E->F is read only by the solver,
 * which finds them on the device */
void __wrap_assemble_forces(struct All_variables *E, int penalty)
{
	if(!ctx || !t_on_device)
	{
		__real_assemble_forces(E, penalty);
		return;
	}
	CHECK(chip_assemble_forces(ctx));
	f_on_device = 1;
}
