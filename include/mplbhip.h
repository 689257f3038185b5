/*
 * mplbhip.h -- C ABI of the MI355X lattice-update library (libmplbhip.so).
 *
 * This is the drop-in boundary under MPLB's HiLeMMS case API.  Each entry
 * point replaces one reference interface (cited per function); the C++
 * HiLeMMS layer (mplb_amd/hilemms) and the Python binding (mplb_amd/hip.py)
 * both sit on top of it.
 *
 * Conventions
 *  - Every call returns 0 on success or a negative code; the message of the
 *    last failure is kept per context (mplbhip_last_error).  No exception
 *    crosses this boundary.  Errors are reported, never silently absorbed:
 *    there is no CPU fallback.
 *  - Plain pointers and sizes only.  Host arrays are owned by the caller;
 *    device memory is owned by the context.
 *  - Host field layout: interior nodes only, i fastest, [k][j][i]; the
 *    distribution function adds the population index outermost,
 *    [q][k][j][i].  2D cases use k = 0 only.
 *  - One context = one case (one component) on one device, driven by one
 *    host thread; calls on one context are not thread-safe.
 *  - Enumerations carry the reference's integer values (VertexType,
 *    BoundarySurface, BoundaryScheme, BodyForceType).
 */
#ifndef MPLBHIP_H
#define MPLBHIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mplbhip_ctx mplbhip_ctx;

/* population storage in HBM */
enum {
    MPLBHIP_STORE_F64 = 0,   /* f in double: the bit-exact parity path */
    MPLBHIP_STORE_F32 = 1,   /* f - w_i in float, FP32 arithmetic */
    MPLBHIP_STORE_F16 = 2    /* (f - w_i)/s_i in half, FP32 arithmetic */
};

enum {
    MPLBHIP_ERR_ARG = -1,     /* bad argument or call order */
    MPLBHIP_ERR_HIP = -2,     /* a HIP runtime call failed */
    MPLBHIP_ERR_UNSUPPORTED = -3,
    MPLBHIP_ERR_NODEVICE = -4 /* compute requested on a plan-only context */
};

/* Create a context.  device >= 0 selects a GPU; device = -1 builds a
 * plan-only context (geometry, node typing and boundary programs can be
 * inspected, nothing is computed).  space_dim 2 selects D2Q9, 3 D3Q19
 * (DefineCase flowfield.cpp:96-104 + DefineComponents model.cpp:208-310). */
int mplbhip_create(mplbhip_ctx** ctx, int device, int space_dim, int storage);
void mplbhip_destroy(mplbhip_ctx* ctx);
const char* mplbhip_last_error(const mplbhip_ctx* ctx);
/* library build string (arch, version) */
const char* mplbhip_version(void);

/* DefineBlocks (flowfield.cpp:160-223): ids 0..n-1 in order; size has
 * space_dim entries, each at least 3. */
int mplbhip_define_block(mplbhip_ctx* ctx, int id, const int* size);
/* DefineBlockConnection (flowfield.cpp:258-283), one pair per call:
 * fStage halos are copied from from_block's from_surface slab to
 * to_block (Field::CreateHalos field.h:215-395). */
int mplbhip_define_connection(mplbhip_ctx* ctx, int from_block,
                              int from_surface, int to_block, int to_surface,
                              int vertex_type);
/* No reference counterpart (the reference joins same-axis faces only,
 * field.h:207-214): a connection between two perpendicular faces, for
 * periodic inlet/outlet pairs of bent channels.  The from-face slab lands in
 * the halo beyond to_surface turned by the quarter rotation about the third
 * axis that takes the from-face's outward normal to the to-face's inward
 * normal; populations are permuted with it.  Transverse sizes must match
 * under the rotation. */
int mplbhip_define_rotated_connection(mplbhip_ctx* ctx, int from_block,
                                      int from_surface, int to_block,
                                      int to_surface, int vertex_type);
/* DefineComponents lattNames (model.cpp:208-310; SetupD2Q9Latt :99-112,
 * SetupD3Q19Latt :114-131, SetupD3Q15Latt :133-150), before mplbhip_prepare:
 * the number of velocities, 9 (2D), 19 (3D, the default) or 15 (3D). */
int mplbhip_define_lattice(mplbhip_ctx* ctx, int q);
/* DefineCollision (model.cpp:406-445), before mplbhip_prepare:
 * 0 Collision_BGKIsothermal2nd (the default), 1 Collision_BGKThermal4th
 * (KerCollideBGKThermal[3D], model_kernel.inc:267-310, 457-502: per node
 * tau = tauRef / (rho sqrt(T)), the 4th-order equilibrium at T and
 * f - w (f - feq) + tau w F; FP64 storage). */
int mplbhip_define_collision_type(mplbhip_ctx* ctx, int type);
/* The Variable_T macro field of one block (interior [k][j][i], T > 0),
 * which the thermal collision reads; as in the reference nothing updates
 * it.  After mplbhip_prepare, before the first step. */
int mplbhip_set_temperature(mplbhip_ctx* ctx, int block, const double* T);
/* DefineScheme (scheme.cpp:92-109), before mplbhip_prepare: 10
 * Scheme_StreamCollision (the default), 11 Scheme_StreamCollision_Swap (the
 * swap scheme of evolution.cpp:195-252: one lattice f updated in place by
 * collision, force, halo exchange of f, local swap, swap stream and the
 * boundaries; D3Q19 and FP64 storage).  FDPeriodic boundaries need the
 * swap scheme: they copy f's halo, which only it exchanges. */
int mplbhip_define_scheme(mplbhip_ctx* ctx, int scheme);
/* No reference counterpart (the reference only skips ImmersedSolid nodes,
 * model_kernel.inc:420, scheme_kernel.inc:1418; its bounce-back is
 * commented out, boundary_kernel.inc:490-532): in-domain obstacles of one
 * block, interior nodes [k][j][i], nonzero = obstacle.  Obstacle nodes are
 * ImmersedSolid; a node pulling from one takes the half-way bounce-back
 * f_q(x) = fs_opp(q)(x) of a resting wall.  Obstacles may sit on Fluid or
 * MDPeriodic nodes whose pulling neighbours pull every population; prepare
 * refuses others.  Before mplbhip_prepare. */
int mplbhip_define_obstacles(mplbhip_ctx* ctx, int block, const unsigned char* solid);
/* DefineComponents tauRef (model.cpp:236) + DefineCollision
 * (Collision_BGKIsothermal2nd only, model.cpp:406-445) + DefineBodyForce
 * (model.cpp:447-489, force_type BodyForce_None=0 / BodyForce_1st=1) +
 * the macro kind of DefineMacroVars (force_macros=1 selects
 * Variable_U_Force..., model_kernel.inc:665-759). */
int mplbhip_define_collision(mplbhip_ctx* ctx, double tau_ref,
                             int force_type, int force_macros);
/* DefineBlockBoundary (boundary.cpp:61-130); the call order is the
 * application order (ImplementBoundary boundary.cpp:379-386).  Schemes:
 * EQMDiffuseRefl (24, vars = wall velocity u,v[,w]), ExtrapolPressure1ST
 * (16, vars = rho; stream-collision scheme), FDPeriodic (19, swap scheme:
 * KerCutCellPeriodic[3D] boundary_kernel.inc:615-719, 1278-1666), None (-1)
 * / MDPeriodic (18) (node typing only). */
int mplbhip_define_boundary(mplbhip_ctx* ctx, int block, int surface,
                            int scheme, const double* vars, int nvars,
                            int vertex_type);
/* Partition -> PrepareFlowField (flowfield.cpp:108-111, 224-244): node
 * typing, geometry, per-region boundary programs, HBM allocation. */
int mplbhip_prepare(mplbhip_ctx* ctx);
/* node type (which=0) or geometry property (which=1) of one block, as the
 * reference's NodeType_<compo> / GeometryProperty fields; which=2 and 3
 * describe the launch plan (no reference counterpart): 1 where the node's
 * boundary program may run inside the bulk kernel, and the faces (bit 2a
 * low, 2a+1 high side of axis a) across which such a node streams nothing. */
int mplbhip_node_info(mplbhip_ctx* ctx, int block, int which, int* out);

/* SetInitialMacrosVars upload (app hook, e.g. lbm3d_cavity.cpp:44-68);
 * w is ignored in 2D and may be NULL there. */
int mplbhip_set_macros(mplbhip_ctx* ctx, int block, const double* rho,
                       const double* u, const double* v, const double* w);
/* PreDefinedInitialCondition[3D] (model_wrapper.cpp:347-391): f = feq of
 * the uploaded macros at every non-solid node. */
int mplbhip_init_feq(mplbhip_ctx* ctx);
/* No reference counterpart (warm start from a predicted flow): f = feq of
 * the uploaded macros plus the first-order Chapman-Enskog non-equilibrium
 * part of BGK, f1_q = -3 tau_L w_q rho (c_qa c_qb - delta_ab/3) d_a u_b in
 * lattice units, with central-difference gradients (one-sided at block
 * faces and obstacles).  Obstacle nodes get f = 0. */
int mplbhip_init_lift(mplbhip_ctx* ctx);
/* restart: set f of every block (array per block, [q][k][j][i]) */
int mplbhip_set_f(mplbhip_ctx* ctx, int block, const double* f);
/* SetTimeStep (flowfield.cpp:146) */
int mplbhip_set_timestep(mplbhip_ctx* ctx, double dt);
/* acceleration of one block (g_MacroBodyforce, uniform per block) */
int mplbhip_set_accel(mplbhip_ctx* ctx, int block, const double* g);

/* StreamCollision x nsteps (evolution.cpp:137-193), queued on the
 * context stream and returned once complete. */
int mplbhip_step(mplbhip_ctx* ctx, int nsteps);
/* nsteps timed with HIP events on the context stream around each step
 * kernel: ms_total is the wall span, ms_kernel the summed time of the step
 * kernels (bulk and boundary; halo copies excluded) */
int mplbhip_step_timed(mplbhip_ctx* ctx, int nsteps, double* ms_total,
                       double* ms_kernel);
/* Diagnostic, no reference counterpart: the bulk-kernel and boundary-kernel
 * shares of the last mplbhip_step_timed call (ms, summed over its steps). */
int mplbhip_kernel_times(mplbhip_ctx* ctx, double* ms_bulk, double* ms_boundary);
/* Diagnostic, no reference counterpart: the step's launch shape in force
 * (after the first step's autotuning): fold level (0 none, 1 simple x faces,
 * 2 every simple region of the planes z = 1 .. n-2, 3 every simple region
 * of every plane, 3D only; MPLBHIP_XFACE_FOLD fixes it), bulk workgroup
 * shape (0 plain tiles, 1 rows dealt to one XCD, 2 row workgroups), 1 if
 * FP16 steps use the two-node kernel, and the nodes one bulk-kernel launch
 * updates (interior nodes plus the folded boundary nodes). */
int mplbhip_launch_shape(mplbhip_ctx* ctx, int* fold, int* xgroup, int* h2,
                         long long* bulk_nodes);
/* Iterate (evolution.cpp:71-135 and the templates evolution.h:78-129).
 * criteria < 0: run exactly max_steps steps, refreshing the macro variables
 * at every check_period-th step (the transient form).  criteria >= 0: step
 * until the residual taken every check_period steps drops below criteria,
 * or max_steps have been taken.  The macros and residual of a check step
 * come out of that step's own kernel.  *steps_done and *residual (the last
 * residual, or -1 if none was taken) may be NULL. */
int mplbhip_iterate(mplbhip_ctx* ctx, double criteria, int check_period,
                    long long max_steps, long long* steps_done,
                    double* residual);
/* ---- independent instances (batched cases) ------------------------------
 * A context may hold many unconnected blocks, each its own case (e.g. a
 * batch of cavities of different sizes and Reynolds numbers): they advance
 * together in one launch per step.  Per-block relaxation time and time step
 * override the context-wide ones (no reference counterpart: the reference
 * runs one case per process). */
int mplbhip_define_block_collision(mplbhip_ctx* ctx, int block, double tau_ref,
                                   double dt);
/* Step every instance until its own velocity residual (the CalcResidualError
 * ratio of u, v[, w] summed over the instance's blocks, maximum over
 * components, divided by check_period * dt) falls below criteria, freezing
 * each instance as it converges (a frozen block costs one early-exit per
 * workgroup per step; its halo copies stop).  An instance is a block, or
 * the blocks joined by connections (a periodic channel's self-linked block,
 * a channel of several linked blocks); its blocks must share one time step.
 * steps_per_block[b] receives the step at which block b's instance
 * converged (-1 if it did not within max_steps); residual_per_block[b] its
 * last residual; *steps_done the steps taken. */
int mplbhip_iterate_batch(mplbhip_ctx* ctx, double criteria, int check_period,
                          long long max_steps, long long* steps_per_block,
                          double* residual_per_block, long long* steps_done);
/* No reference counterpart: the batched solver's fallback for warm-started
 * instances.  With factor > 0, an instance whose velocity residual at a
 * check exceeds factor times the lowest it has reached is restarted once at
 * rest (both fStage buffers = feq(1, 0), macros rho = 1, u = 0) and goes on
 * from there.  0 (the default) turns it off.  Refused for cases whose
 * boundaries persist state (3D corners). */
int mplbhip_set_batch_restart(mplbhip_ctx* ctx, double factor);
/* restarts per block during the last mplbhip_iterate_batch */
int mplbhip_batch_restarts(mplbhip_ctx* ctx, int* restarts);

/* UpdateMacroVars[3D] (model_wrapper.cpp:112-275): macros of the current f */
int mplbhip_update_macros(mplbhip_ctx* ctx);
/* current distribution function (post boundary conditions), [q][k][j][i] */
int mplbhip_get_f(mplbhip_ctx* ctx, int block, double* out);
/* macro variable 0=rho 1=u 2=v 3=w as last updated */
int mplbhip_get_macro(mplbhip_ctx* ctx, int block, int var, double* out);
/* CalcResidualError + GetMaximumResidual (flowfield_wrapper.cpp:31-89,
 * flowfield.cpp:148-158) on the current f, in FP64 on the device.
 * per_var receives, for rho, u, v[, w], the ratio
 * sum (m - m_copy)^2 / sum m^2 that CalcResidualError stores; *max is
 * GetMaximumResidual(check_period) = max ratio / (check_period * dt).
 * MacroVarsCopy is refreshed, as in the reference. */
int mplbhip_residual(mplbhip_ctx* ctx, int check_period, double* per_var,
                     double* max);
/* the ratios and maximum of the last residual taken, by mplbhip_residual
 * or by a check step of mplbhip_iterate */
int mplbhip_last_residual(mplbhip_ctx* ctx, double* per_var, double* max);
/* FP64 total mass of the current f over every interior node */
int mplbhip_mass(mplbhip_ctx* ctx, double* mass);
/* No counterpart in the reference: mplbhip_mass for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops): it calls no hipMalloc or hipFree.  Its partial sums live in a
 * file-scope device array of 1 Mi doubles (8 MiB) per GPU, one slot per
 * workgroup of 256 nodes of the launches that measure it (the same as
 * mplbhip_mass's), so a case may have at most about 116,000 such
 * workgroups; larger cases are refused with
 * MPLBHIP_ERR_UNSUPPORTED.  Needs at least one step taken (f^0 is formed on
 * the device by the first step).  Calls on one GPU share the array: calls
 * that overlap there race.  Same value as mplbhip_mass, bit for bit. */
int mplbhip_mass_noalloc(mplbhip_ctx* ctx, double* mass);
/* steps taken since the initial condition */
int mplbhip_time_step(mplbhip_ctx* ctx, long long* step);
int mplbhip_sync(mplbhip_ctx* ctx);

/* ---- diagnostics (no reference counterpart) ----------------------------
 * Plain device-to-device copy of nbytes (16-byte vector loads and stores)
 * repeated iters times on `device`; ms_per_iter is the HIP-event average.
 * Used to calibrate HBM traffic counters and as the achievable-bandwidth
 * yardstick next to the lattice kernel. */
int mplbhip_copy_probe(int device, long long nbytes, int iters,
                       double* ms_per_iter);

#ifdef __cplusplus
}
#endif
#endif
