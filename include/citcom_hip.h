/*
 * citcom_hip.h -- C ABI of libcitcom_hip.so, the MI355X (gfx950) implementation
 * of CitcomCU's timestep hot path: the Uzawa pressure CG around the full-
 * multigrid velocity solve, and the Petrov-Galerkin (SUPG) energy step.
 *
 * Plain C types only.  Every entry point returns 0 on success and a negative
 * code on failure; chip_last_error() then describes the failure.  Host arrays
 * use the reference's numbering made 0-based:
 *   node  n = k + noz*(i + nox*j)    (k: z index, i: x index, j: y index,
 *                                     Construct_arrays.c:188, Nodal_mesh.c:182)
 *   element e = ez + elz*(ex + elx*ey)        (Construct_arrays.c:70-72)
 *   velocity dof 3n+d, d = 0:x 1:y 2:z        (ID doff, Construct_arrays.c:158-165)
 * i.e. E->U[0..neq-1], E->P[1..npno], E->T[1..nno], E->V[d+1][1..nno] of the
 * reference map to these arrays with the 1-based offset removed.
 */
#ifndef CITCOM_HIP_H
#define CITCOM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Problem description: the control data of struct All_variables that the hot
 * path reads (global_defs.h:931-1102), set by read_instructions
 * (Instructions.c:55-157) in the reference. */
typedef struct chip_problem
{
	int levels;	/* levmax = levels-1, levmin = 0 (Instructions.c:593-598) */
	int mgunitx, mgunity, mgunitz;
	const float *xc, *yc, *zc;	/* finest-level node coordinates per axis */
	int topvbc, botvbc;	/* 0: free slip, 1: prescribed (zero) velocity */
	int toptbc, bottbc;	/* 1: fixed temperature */
	float toptbcval, bottbcval;
	double accuracy;	/* control.accuracy */
	int p_iterations;	/* piterations */
	int v_steps_low, v_steps_high, down_heavy, up_heavy, mg_cycle;
	int aug_lagr;
	double aug_number;
	int precond;
	int rheol, tdepv, num_mat;	/* viscosity law, Viscosity_structures.c:491-737 */
	int visc_update;	/* VISC_UPDATE: rebuild viscosity and stiffness every step */
	float visc0[4], viscE[4], viscT[4], z_layer[4];
	int vmin_on, vmax_on;
	float visc_min, visc_max;
	float rayleigh, finetunedt;
	int temp_iterations;	/* adv_sub_iterations */
	float gamma;
	double Q0;
	float fixed_timestep;
	int mixed;	/* 0: reference arithmetic (FP64 vectors, FP32 stored K);
			   1: FP32 multigrid inside FP64 iterative refinement;
			   2: as 1, with the two finest levels' stiffness rows scaled
			      by BI and stored in FP16 for the multigrid cycles */
} chip_problem;

typedef struct chip_ctx chip_ctx;

/* Uzawa monitors written by solve_Ahat_p_fhat (Stokes_flow_Incomp.c:370-446) */
typedef struct chip_stokes_stats
{
	int iterations;	/* Uzawa iterations (count) */
	int mg_cycles;	/* multi_grid calls over all inner solves */
	int inner_solves;	/* solve_del2_u calls */
	double vdotv, pdotp, incompressibility;
	double dvelocity, dpressure;
} chip_stokes_stats;

typedef struct chip_pg_stats
{
	float timestep;	/* advection.timestep */
	float T_interior;	/* monitor.T_interior (Tmax after the step) */
	int sub_iterations;	/* advection.last_sub_iterations of the retry loop */
	float elapsed_time;	/* monitor.elapsed_time */
} chip_pg_stats;

const char *chip_last_error(void);

/* Builds the mesh of every multigrid level on the current HIP device:
 * coordinates, boundary flags, Jacobians and shape-function derivatives,
 * lumped mass and nodal weights.  Replaces the mesh part of
 * read_instructions (Instructions.c:94-136): construct_ien/masks/id
 * (Construct_arrays.c:52-211), node_locations (Nodal_mesh.c:53-283),
 * velocity/temperature_boundary_conditions (Boundary_conditions.c:44-246),
 * construct_shape_functions (Shape_functions.c:53-170), mass_matrix
 * (Size_does_matter.c:618-757). */
int chip_create(const chip_problem *p, chip_ctx **out);
void chip_destroy(chip_ctx *c);

/* Stream all work is queued on (a hipStream_t; NULL = the library's own). */
int chip_set_stream(chip_ctx *c, void *hip_stream);
void *chip_get_stream(chip_ctx *c);

/* ---- state (host <-> device copies; sizes in elements) ------------------- */
/* which: "T" float[nno], "Tdot" float[nno], "U" double[3nno], "P" double[nel],
 * "F" double[3nno], "V" float[3*nno] (V[d*nno+n]), "buoyancy" float[nno],
 * "EVI" float[8*nel(lev)] (lev = -1: finest), "K" float[nno(lev)*243] row
 * layout [n][r*81 + slot*3 + c], slot = (dz+1)+3(dx+1)+9(dy+1),
 * "BI" double[3nno(lev)], "BPI" double[nel(lev)] */
int chip_upload(chip_ctx *c, const char *which, int lev, const void *host, long count);
int chip_download(chip_ctx *c, const char *which, int lev, void *host, long count);

/* ---- operator build (stiffness rebuild) ---------------------------------- */
/* get_system_viscosity (Viscosity_structures.c:369-443) from the device T:
 * visc_from_T / visc_from_mat and the MIN/MAX clip, into EVI[levmax]. */
int chip_get_system_viscosity(chip_ctx *c);
/* construct_stiffness_B_matrix (Construct_arrays.c:834-889): project_viscosity
 * (Solver_multigrid.c:398-480), construct_elt_gs, construct_node_ks with
 * get_elt_k/get_aug_k/build_diagonal_of_K (Construct_arrays.c:359-535,
 * Element_calculations.c:133-293, 624-652, 1127-1163), build_diagonal_of_Ahat
 * (Element_calculations.c:654-685), from EVI[levmax]. */
int chip_construct_stiffness_B_matrix(chip_ctx *c);

/* ---- single operators, for verification (host arrays in and out) --------- */
/* op: "matvec" n_assemble_del2_u (Element_calculations.c:552-621, strip on),
 *     "gs0"/"gs1" gauss_seidel guess 0/1 with arg sweeps, 8-colour
 *         (General_matrix_functions.c:1160-1361); in = d0 (gs1), in2 = F,
 *         out = d0 after, out2 = A d0
 *     "box0"/"box1" (levels with the box smoother of the FP32 cycles, mixed
 *         forms) arg sweeps of the box smoother in FP32 from zero / from in,
 *         then A d0: in, in2, out, out2 as for "gs0"/"gs1"
 *     "coarse" (coarsest level) the multigrid's coarsest-level smoother call
 *         (v_steps_low sweeps from zero) applied as its matrix: out = Mc in2
 *     "affine0"/"affine1" (the level above) its multigrid smoother call
 *         (down_heavy sweeps) as an affine map: out = Mb in2 (zero guess) or
 *         Ma in + Mb in2
 *     "cnet" (coarsest level) the learned coarse-level correction for the
 *         right-hand side in2 (after chip_set_coarse_net)
 *     "direct" (the level the FP32 cycles solve exactly, mixed forms, after
 *         a rebuild) the FP32 inverse of its stiffness: out = K^-1 in2
 *     "project" project_vector lev -> lev-1 (Solver_multigrid.c:72-159)
 *     "interp"  interp_vector lev-1 -> lev (Solver_multigrid.c:173-298)
 *     "div" assemble_div_u (Element_calculations.c:691-720)
 *     "grad" assemble_grad_p (Element_calculations.c:727-769)
 *     "strip" strip_bcs_from_residual (Boundary_conditions.c:926-947) */
int chip_apply(chip_ctx *c, const char *op, int lev, const double *in, const double *in2, double *out, double *out2, int arg);

/* ---- drop-in #1: solve_constrained_flow_iterative ------------------------ */
/* Stokes_flow_Incomp.c:50-90: Uzawa CG (solve_Ahat_p_fhat :295-497) on the
 * device U, P with right-hand side F, then v_from_vector (:530-552) into V.
 * With the option "warm_start" the device U, P are first replaced by the
 * warm start's guess (and the velocity boundary values imposed again). */
int chip_solve_constrained_flow_iterative(chip_ctx *c, chip_stokes_stats *st);
/* solve_del2_u (General_matrix_functions.c:368-520) on host arrays, for
 * verification: d0 = K^-1 F to max(acc, accuracy*|F|) */
int chip_solve_del2_u(chip_ctx *c, const double *F, double *d0, double acc, int *valid, int *cycles);

/* ---- drop-in #2: PG_timestep --------------------------------------------- */
/* Advection_diffusion.c:251-345 with std_timestep, predictor, pg_solver,
 * corrector, Tmax retry, temperatures_conform_bcs, thermal_buoyancy. */
int chip_PG_timestep(chip_ctx *c, chip_pg_stats *st);

/* ---- per-step inputs of the Stokes solve ---------------------------------- */
/* thermal_buoyancy (Pan_problem_misc_functions.c:83-153) from the device T */
int chip_thermal_buoyancy(chip_ctx *c);
/* assemble_forces (Element_calculations.c:74-125) from the device buoyancy */
int chip_assemble_forces(chip_ctx *c);
/* heat_flux (Process_buoyancy.c:63-203): nu[0] = Nu_top, nu[1] = Nu_bot */
int chip_heat_flux(chip_ctx *c, float *nu);
/* volume-RMS velocity of the device V: sqrt(int v.v dV / int dV) with the
 * 2x2x2 Gauss quadrature of the reference's shape functions and Jacobians
 * (Shape_functions.c:53-170, Size_does_matter.c:60-179); the reference has
 * no routine for it */
int chip_vrms(chip_ctx *c, double *vrms);

/* averages (Process_velocity.c:179-224) and the temperature profile of
 * heat_flux (Process_buoyancy.c:93): per node layer k = 0..noz-1 the
 * horizontal averages (return_horiz_ave, Global_operations.c:133-251) of
 * T (Have.T), of |V|^2 square-rooted (Have.vrms) and of the nodal
 * viscosity of the finest EVI (visc_from_gint_to_nodes, Nodal_mesh.c:583-615;
 * Have.Vi), from the device T, V and EVI.  Host arrays of noz floats; any
 * may be NULL. */
int chip_averages(chip_ctx *c, float *T_ave, float *vrms_ave, float *visc_ave);

/* One timestep of the reference's time loop (Citcom.c:111-161) with every
 * stage on the device: PG_timestep, heat_flux, then general_stokes_solver
 * (Drive_solvers.c:45-165: velocities_conform_bcs, assemble_forces, viscosity
 * and stiffness update when the viscosity is temperature dependent, Uzawa). */
int chip_timestep(chip_ctx *c, chip_pg_stats *pg, chip_stokes_stats *st, float *nu);
/* general_stokes_solver alone (the step-0 solve of Citcom.c:97) */
int chip_general_stokes_solver(chip_ctx *c, chip_stokes_stats *st);

/* ---- measurement ---------------------------------------------------------- */
/* kernel: "gs_colour", "matvec", "gs_wg" (the one-workgroup smoother call
 * of small levels), "gs_xcd" (the one-XCD smoother call of mid-size
 * levels) or "gs_box" (one sweep of the box smoother of the FP32
 * cycles).  While enabled, every launch of that kernel
 * is bracketed by HIP events on the work stream; read returns the number of
 * launches, their summed duration and the algorithmic bytes they moved. */
int chip_profile_enable(chip_ctx *c, const char *kernel, int on);
int chip_profile_read(chip_ctx *c, long *launches, double *total_ms, double *bytes);
/* device memory held by the context, bytes */
long chip_device_bytes(chip_ctx *c);
/* "graphs": 1 (default) replays each multigrid call as one captured HIP
 * graph, 0 launches its kernels one by one; results are identical.
 * "small_levels": 1 (default) smooths levels whose stiffness fits in LDS in
 * one single-workgroup kernel per gauss_seidel call, 0 launches one kernel
 * per colour pass on every level; results are identical.
 * "gs_wg": 1 (default) also runs a whole gauss_seidel call in one workgroup
 * on levels of at most 1024 nodes whose stiffness does not fit in LDS;
 * 0 launches one kernel per colour pass there; results are identical.
 * "coarse_direct": 1 (default) lets the FP32 cycles of the mixed forms
 * stop at the largest level below the finest with at most 8192 unknowns
 * and at least two levels below it (17x17x9 nodes at 128x128x64 elements)
 * and solve it exactly: its stiffness is assembled dense, factored and
 * inverted in FP64 by rocSOLVER (potrf, potri) at every rebuild and
 * applied as FP16 rows scaled by their largest entry (FP64 sums), in
 * place of the smoother calls, transfers and maps of
 * that level and the ones below; 0 keeps the full hierarchy.  The plain
 * form always keeps it.  The solve still iterates to the same criterion.
 * "direct_async": 1 (default) factors every rebuild's matrix after the
 * first on a side stream while the step's solve goes on with the previous
 * inverse (taken up at the next rebuild); 0 factors on the work stream and
 * uses the new inverse at once.
 * "xcd_gs": 1 runs a whole gauss_seidel call on levels of 1025 to 6000
 * nodes as one launch of a few workgroups on one XCD, with barriers through
 * that XCD's L2 between the colour passes; 0 (default) launches one kernel
 * per colour pass there; results are identical.  Setting 1 first checks
 * that the device dispatches workgroups to its XCDs round-robin, and is
 * refused where it does not.
 * "box_gs": 1 (default) smooths the levels of the mixed forms' FP32 cycles
 * between the exact (or the affine) level and the finest, larger than 1024
 * nodes, by boxes of at most 8 nodes a side: one launch per sweep, a
 * workgroup per box running the 8 colour passes on its own nodes in LDS,
 * the other boxes' nodes as they were before the sweep; 0 launches one
 * kernel per colour pass there.  The iterates differ (a box's face sees its
 * neighbour's values of the previous sweep); the solve still iterates to
 * the same criterion.
 * "profile_level": >= 0 restricts chip_profile_enable to that level's
 * launches, -1 (default) profiles every level.
 * "coarse_net": 1 replaces the coarsest level's smoother call in multigrid
 * by the learned correction (chip_set_coarse_net), 0 (default) does not.
 * "warm_start": 0 (default) starts each Uzawa solve from the previous
 * solution as the last solve left it, as the reference does
 * (Stokes_flow_Incomp.c:354-366 start from E->U, E->P); 1 starts it from
 * the last solutions extrapolated along the sequence of solves: the
 * velocity 2 U_n - U_(n-1), the pressure with its increment averaged over
 * the last two solves 3:1, P_n + 3/4 (P_n - P_(n-1)) + 1/4 (P_(n-1) -
 * P_(n-2)) (linearly, 2 P_n - P_(n-1), while only two solves are kept); 2
 * adds the learned U-Net's correction of that extrapolation
 * (chip_set_warm_net) once three have been kept.  Setting the option
 * starts a new history.  The solve still iterates to the same criterion.
 * "print_convergence": 1 prints one line per Uzawa iteration and per inner
 * solve on stderr (generate_log_message, Stokes_flow_Incomp.c:501-526;
 * General_matrix_functions.c:494-498). */
int chip_set_option(chip_ctx *c, const char *name, int value);

/* ---- learned coarse-level correction (no counterpart in the reference) ---- */
/* A 3-D CNN on the coarsest multigrid level (BF16 MFMA convolutions, FP32
 * accumulation; the kernels of citcom_learn.h) that takes the place of the
 * coarsest level's smoother call in multi_grid (General_matrix_functions.c
 * :574, :613) once the option "coarse_net" is 1.  Node inputs: the
 * right-hand side over its RMS (3 channels), log10 viscosity (1 channel);
 * output channels 0..2 times out_scale[0..2] times that RMS, boundary rows
 * zero.  weights: BF16 bits [layers][16 out][27 taps][16 in], bias
 * [layers][16]; layers 1..8, ReLU after all but the last; a coarsest level
 * of at most 1024 nodes (one workgroup runs the network).  Copied to the
 * device; the inner solve keeps iterating to its own residual criterion. */
int chip_set_coarse_net(chip_ctx *c, int layers, const uint16_t *weights, const float *bias,
                        const float *out_scale);
/* The plain multigrid is its fallback.  The first inner solve
 * (solve_del2_u) after the option "coarse_net" is set to 1 runs with the
 * coarsest level's smoother and again with the correction; the correction
 * stays on only if its mean residual reduction per multigrid call is the
 * stronger, else the plain solve's result is kept and the correction is
 * switched off.  Later, an inner solve whose multigrid call with the
 * correction does not reduce the residual switches it off for the rest of
 * the context's life and restarts with the coarsest level's smoother.  on:
 * whether it is still in use; fallbacks: how many times it was switched
 * off. */
int chip_coarse_net_state(chip_ctx *c, int *on, int *fallbacks);

/* ---- learned warm start (no counterpart in the reference) ----------------
 * The Uzawa solve's initial guess with the option "warm_start" = 2: the last
 * solutions extrapolated along the sequence of solves (as with
 * "warm_start" = 1), corrected by a
 * U-Net (BF16 MFMA convolutions, FP32 accumulation; cl_unet_forward of
 * citcom_learn.h) on the finest level's nodes.  Its input per node, 16 bf16
 * channels: T - 1/2; dT / s_dT with dT = T - (2 T_n - T_(n-1)) (T_n,
 * T_(n-1) the temperatures of the last two solves); log10(eta) / 2.5
 * (Gauss-point mean of the adjacent elements, averaged); the extrapolated
 * U / s_u (3) and its nodal P / s_p (mean of the adjacent elements); the
 * last increments U_n - U_(n-1) / s_u (3) and of the nodal P / s_p; the
 * second differences U_n - 2 U_(n-1) + U_(n-2) / s_2u (3) and of the nodal
 * P / s_2p; z (node index over noz - 1); s_* the RMS values.  Its output
 * channels 0..2 times s_2u are added to U at every node and channel 3 times
 * s_2p, averaged over an element's eight nodes, to its P.  The solve then
 * runs to its own criterion, so accuracy is not traded.  Weights: bf16
 * [594432] in layer order ([co][27][ci] each), biases FP32 [528]. */
int chip_set_warm_net(chip_ctx *c, const uint16_t *weights, long n_weights, const float *bias, long n_bias);
/* Training samples: with the warm start on, every solve that has three
 * earlier ones writes the network's input (bf16 [nno][16]) to features and, after
 * the solve, the target (FP32 [nno][4]: (U - U_extrapolated) / s_2u per
 * node, the nodal mean of (P - P_extrapolated) / s_2p) to targets; device
 * pointers, both NULL to stop.  Samples start once three solves are kept. */
int chip_warm_record(chip_ctx *c, void *features, void *targets);
/* the last formed input's scales: s_u, s_p, s_dT, s_2u, s_2p, 0, the
 * number of solves kept before this one, 1 (0 before any input formed) */
int chip_warm_scales(chip_ctx *c, double *out8);

#ifdef __cplusplus
}
#endif
#endif
