/*
 * mplb_oracle.h -- CPU oracle for the MPLB per-step lattice update.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, sequentially-ordered
 * restatement of the reference algorithm (inmeso/mplb, Src/), used as the
 * checker for the HIP product path and as the timed CPU baseline.  Only the
 * tests, the smoke entry point and the benchmark's cpu_baseline leg may load
 * it; the product library never links, calls or falls back to it.
 *
 * Every function names the reference file:line it restates.  The data model
 * mirrors OPS dats: one array per block and field, halo depth 1 on every
 * axis (no z halo in 2D), distribution components innermost (AoS), the
 * i index fastest.
 */
#ifndef MPLB_ORACLE_H
#define MPLB_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_case orc_case;

/* lattice: 0 = D2Q9 (model.cpp:99-112), 1 = D3Q19 (model.cpp:114-131),
 * 2 = D3Q15 (model.cpp:133-150) */
orc_case* orc_create(int space_dim, int lattice);
void orc_destroy(orc_case* c);
const char* orc_error(orc_case* c);

/* DefineBlocks (flowfield.cpp:160-223); blocks must be added in id order. */
int orc_add_block(orc_case* c, int id, int nx, int ny, int nz);
/* DefineBlockConnection (flowfield.cpp:258-283) */
int orc_add_connection(orc_case* c, int from_block, int from_surface,
                       int to_block, int to_surface, int vertex_type);
/* Extension with no reference counterpart: a connection between two
 * perpendicular faces, turned by the quarter rotation about the third axis
 * that takes the from-face's outward normal to the to-face's inward one. */
int orc_add_rotated_connection(orc_case* c, int from_block, int from_surface,
                               int to_block, int to_surface, int vertex_type);
/* Extension with no reference counterpart: obstacle nodes of a block
 * (interior [k][j][i], nonzero = solid), at least two nodes inside every
 * face; Fluid neighbours bounce back half-way.  Before orc_prepare. */
int orc_set_obstacles(orc_case* c, int block, const unsigned char* solid);
/* DefineComponents tauRef + DefineBodyForce (model.cpp:208-310, 447-489).
 * force_type: 0 = BodyForce_None, 1 = BodyForce_1st.
 * force_macros: 1 selects the Variable_*_Force velocity kernels
 * (model_kernel.inc:665-759), 0 the plain ones (model_kernel.inc:593-663). */
int orc_set_collision(orc_case* c, double tau_ref, int force_type,
                      int force_macros);
/* DefineCollision (model.cpp:406-445): 0 = Collision_BGKIsothermal2nd,
 * 1 = Collision_BGKThermal4th (model_kernel.inc:267-310, 457-502) */
int orc_set_collision_type(orc_case* c, int type);
/* Variable_T of one block, interior [k][j][i] (the thermal collision reads
 * it; nothing updates it, as in the reference); 1 until set */
int orc_set_temperature(orc_case* c, int block, const double* T);
/* CalcBGKFeq with temperature T and polyOrder (model_host_device.h:34-83) */
double orc_feq_t(orc_case* c, int l, double rho, double u, double v, double w,
                 double T, int poly_order);
/* DefineBlockBoundary (boundary.cpp:61-130); call order is kept. */
int orc_add_boundary(orc_case* c, int block, int surface, int scheme,
                     const double* vars, int nvars, int vertex_type);
/* Partition -> PrepareFlowField (flowfield.cpp:108-111, 224-244) */
int orc_prepare(orc_case* c);
/* App hook SetInitialMacrosVars: interior arrays [k][j][i]; w ignored in 2D */
int orc_set_macros(orc_case* c, int block, const double* rho,
                   const double* u, const double* v, const double* w);
/* PreDefinedInitialCondition[3D] (model_wrapper.cpp:347-391, 641-683) */
int orc_init_feq(orc_case* c);
/* SetTimeStep (flowfield.cpp:146) */
void orc_set_dt(orc_case* c, double dt);
/* g_MacroBodyforce: uniform acceleration of one block (dim values) */
int orc_set_accel(orc_case* c, int block, const double* g);
/* StreamCollision x n (evolution.cpp:137-193) */
int orc_step(orc_case* c, int nsteps);
/* test hook: the boundary stage of a step alone (ImplementBoundary on the
 * current f, every BlockBoundary in definition order) */
int orc_apply_boundaries(orc_case* c);
/* SchemeType: 10 Scheme_StreamCollision (default), 11 Scheme_StreamCollision_Swap */
int orc_set_scheme(orc_case* c, int scheme);
/* UpdateMacroVars[3D] (model_wrapper.cpp:112-275, 468-588) */
void orc_update_macros(orc_case* c);
/* distribution function of one block, interior only, layout [q][k][j][i] */
int orc_get_f(orc_case* c, int block, double* out);
int orc_set_f(orc_case* c, int block, const double* in);
/* macro variable 0=rho 1=u 2=v 3=w, interior [k][j][i] */
int orc_get_macro(orc_case* c, int block, int var, double* out);
/* which: 0 = node type, 1 = geometry property; interior [k][j][i] */
int orc_get_int(orc_case* c, int block, int which, int* out);
/* total mass: sum over blocks, interior nodes, populations of f */
double orc_mass(orc_case* c);
/* CalcResidualError + GetMaximumResidual (flowfield_wrapper.cpp:31-89,
 * flowfield.cpp:148-158).  per_var receives the stored ratios
 * err_v = sum (m - m_copy)^2 / sum m^2 for rho, u, v[, w]; the return value
 * is max_v err_v/(check_period*dt). */
double orc_residual(orc_case* c, int check_period, double* per_var);
/* OpenMP threads used by the sweeps (0 = library default) */
void orc_set_threads(int nthreads);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
