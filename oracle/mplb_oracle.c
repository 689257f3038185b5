/*
 * mplb_oracle.c -- CPU oracle (FP64) for MPLB's per-step lattice update.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the
 * timed CPU baseline.  Nothing in the product library links or calls it.
 *
 * It restates, sweep by sweep and in the reference's floating-point
 * operation order, the StreamCollision step of inmeso/mplb
 * (Src/evolution.cpp:137-193):
 *   macros -> envelope copy -> body force -> BGK collision -> halo transfer
 *   -> pull stream -> block boundaries in definition order.
 * Build with -ffp-contract=off so that no multiply-add is fused: the
 * reference's arithmetic is then reproduced operation for operation.
 *
 * Parity pins (checked by tests/test_oracle_*.py): lattice identities,
 * equilibrium moments, collision invariants, periodic permutation and mass
 * conservation, the rest state as a fixed point of the wall scheme, the
 * shear-wave decay rate, body-force Poiseuille flow, the Ghia et al.
 * cavity profile, and the reference quirks (frozen 3D corner populations,
 * last-definition-wins at shared edges, the extrapolation rescale that
 * skips the last population).  Extensions with no reference counterpart
 * are pinned too: the rotated connection by rotation equivariance
 * (tests/test_rotated_link.py) and the obstacle bounce-back by exact
 * Poiseuille flow at tau = 1/2 + sqrt(3/16) and exact mass conservation
 * (tests/test_obstacles.py).  The halo slab extents at multi-block
 * junction corners, the reduction-handle reset between the two residual
 * sums, and FDPeriodic are "parity unpinned" (no fixture and no closed form
 * reaches them); FDPeriodic is not restated.
 */
#include "mplb_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- enums, values as in the reference headers ---------------------- */
/* VertexType  (Src/flowfield_host_device.h:21-37) */
enum { VT_FLUID = 10, VT_INLET = 11, VT_OUTLET = 12, VT_MDPERIODIC = 13,
       VT_SYMMETRY = 14, VT_FDPERIODIC = 15, VT_WALL = 1000,
       VT_IMMERSEDSOLID = -1, VT_IMMERSEDBOUNDARY = -2,
       VT_VIRTUALBOUNDARY = -3 };
/* VertexGeometryType faces (Src/flowfield_host_device.h:77-144) */
enum { VG_FLUID = 0, VG_IMMERSEDSOLID = -1, VG_IP = 10, VG_IM = 11,
       VG_JP = 20, VG_JM = 21, VG_KP = 30, VG_KM = 31 };
/* BoundarySurface (Src/boundary_host_device.h:7-29) */
enum { BS_LEFT = 0, BS_RIGHT = 1, BS_TOP = 2, BS_BOTTOM = 3, BS_FRONT = 4,
       BS_BACK = 5, BS_LEFTBACK = 50, BS_LEFTFRONT = 40, BS_RIGHTBACK = 51,
       BS_RIGHTFRONT = 41, BS_TOPBACK = 52, BS_TOPFRONT = 42,
       BS_BOTTOMBACK = 53, BS_BOTTOMFRONT = 43, BS_LEFTTOP = 20,
       BS_LEFTBOTTOM = 30, BS_RIGHTTOP = 21, BS_RIGHTBOTTOM = 31 };
/* BoundaryScheme (Src/boundary.h:71-84) */
enum { SCH_EXTRAPOLPRESSURE1ST = 16, SCH_MDPERIODIC = 18,
       SCH_FDPERIODIC = 19, SCH_BOUNCEBACK = 20, SCH_EQMDIFFUSEREFL = 24,
       SCH_NONE = -1 };
/* CollisionType (Src/model.h:94-99) */
enum { COLL_ISOTHERMAL2ND = 0, COLL_THERMAL4TH = 1 };
/* SchemeType (Src/scheme.h:69-72) */
enum { SCHEME_STREAM_COLLISION = 10, SCHEME_SWAP = 11 };
/* BndryDvType (Src/boundary_host_device.h:34-39) */
enum { DV_INCOMING = 1, DV_OUTGOING = 2, DV_PARALLEL = 3, DV_INVALID = -1 };

#define MAXQ 19
#define MAXB 64
#define MAXBC 256
#define MAXLINK 64

typedef struct {
    int id, n[3];      /* interior size; n[2] = 1 in 2D */
    int h[3];          /* halo depth per axis: 1,1,1 (3D) or 1,1,0 (2D) */
    int s[3];          /* allocated extent per axis (n + 2h) */
    size_t nalloc;     /* nodes incl. halo */
    double *f, *fs;    /* AoS [node][q] */
    double *m[4];      /* rho, u, v, w */
    double *mc[4];     /* copies for the residual (MacroVarsCopy) */
    int *nt, *vg;      /* node type, geometry property */
    unsigned char* solid; /* obstacle nodes (extension), NULL when none */
    double g[3];       /* uniform acceleration */
    double* temp;      /* Variable_T (the thermal collision reads it) */
} OBlock;

typedef struct { int from, fsurf, to, tsurf, type, rot; } OLink;
typedef struct { int block, surface, scheme, vtype, nvars;
                 double vars[8]; } OBC;

struct orc_case {
    int dim, Q;
    double CS;
    double XI[MAXQ * 3];   /* stored like the reference: XI[l*D + d] */
    double W[MAXQ];
    int OPP[MAXQ];
    int nb;
    OBlock b[MAXB];
    int nlink;
    OLink link[MAXLINK];
    int nbc;
    OBC bc[MAXBC];
    double tau, dt;
    int force_type, force_macros;
    int scheme;            /* SCHEME_STREAM_COLLISION or SCHEME_SWAP */
    int collision;         /* COLL_ISOTHERMAL2ND or COLL_THERMAL4TH */
    int prepared;
    char err[256];
};

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}
#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for collapse(2) schedule(static) num_threads(orc_get_threads())")
#else
#define OMP_FOR
#endif

static int fail(orc_case* c, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof c->err, fmt, ap);
    va_end(ap);
    return -1;
}
const char* orc_error(orc_case* c) { return c->err; }

/* ---- lattice tables: model.cpp:99-131, FindReverseXi model.cpp:82-97 -- */
static void setup_lattice(orc_case* c, int lattice) {
    c->CS = sqrt(3.0); /* model.cpp:67-68 */
    if (lattice == 0) {
        const double t00 = 4.0 / 9.0, t01 = 1.0 / 9.0, t11 = 1.0 / 36.0;
        const double t[9] = {t00, t11, t01, t11, t01, t11, t01, t11, t01};
        const int cx[9] = {0, -1, -1, -1, 0, 1, 1, 1, 0};
        const int cy[9] = {0, 1, 0, -1, -1, -1, 0, 1, 1};
        const int op[9] = {0, 5, 6, 7, 8, 1, 2, 3, 4};
        c->Q = 9;
        for (int l = 0; l < 9; l++) {
            c->XI[l * 2] = cx[l];
            c->XI[l * 2 + 1] = cy[l];
            c->W[l] = t[l];
            c->OPP[l] = op[l];
        }
    } else if (lattice == 2) {
        /* SetupD3Q15Latt, model.cpp:133-150 */
        const double t000 = 2 / ((double)9), t001 = 1 / ((double)9),
                     t111 = 1 / ((double)72);
        const double t[15] = {t000, t001, t001, t001, t111, t111, t111, t111,
                              t001, t001, t001, t111, t111, t111, t111};
        const int cx[15] = {0, -1, 0, 0, -1, -1, -1, -1, 1, 0, 0, 1, 1, 1, 1};
        const int cy[15] = {0, 0, -1, 0, -1, -1, 1, 1, 0, 1, 0, 1, 1, -1, -1};
        const int cz[15] = {0, 0, 0, -1, -1, 1, -1, 1, 0, 0, 1, 1, -1, 1, -1};
        c->Q = 15;
        for (int l = 0; l < 15; l++) {
            c->XI[l * 3] = cx[l];
            c->XI[l * 3 + 1] = cy[l];
            c->XI[l * 3 + 2] = cz[l];
            c->W[l] = t[l];
        }
        for (int i = 0; i < 15; i++)
            for (int j = 0; j < 15; j++) {
                int rev = 1;
                for (int k = 0; k < 3; k++)
                    rev = rev && (c->XI[i * 3 + k] + c->XI[j * 3 + k] == 0.0);
                if (rev) { c->OPP[i] = j; break; }
            }
    } else {
        const double t000 = 1 / ((double)3), t001 = 1 / ((double)18),
                     t011 = 1 / ((double)36);
        const double t[19] = {t000, t001, t001, t001, t011, t011, t011,
                              t011, t011, t011, t001, t001, t001, t011,
                              t011, t011, t011, t011, t011};
        const int cx[19] = {0, -1, 0, 0, -1, -1, -1, -1, 0, 0,
                            1, 0, 0, 1, 1, 1, 1, 0, 0};
        const int cy[19] = {0, 0, -1, 0, -1, 1, 0, 0, -1, -1,
                            0, 1, 0, 1, -1, 0, 0, 1, 1};
        const int cz[19] = {0, 0, 0, -1, 0, 0, -1, 1, -1, 1,
                            0, 0, 1, 0, 0, 1, -1, 1, -1};
        c->Q = 19;
        for (int l = 0; l < 19; l++) {
            c->XI[l * 3] = cx[l];
            c->XI[l * 3 + 1] = cy[l];
            c->XI[l * 3 + 2] = cz[l];
            c->W[l] = t[l];
        }
        /* brute-force opposite search, first match wins */
        for (int i = 0; i < 19; i++)
            for (int j = 0; j < 19; j++) {
                int rev = 1;
                for (int k = 0; k < 3; k++)
                    rev = rev && (c->XI[i * 3 + k] + c->XI[j * 3 + k] == 0.0);
                if (rev) { c->OPP[i] = j; break; }
            }
    }
}

/* ---- index helpers ---------------------------------------------------- */
static inline size_t nidx(const OBlock* b, int i, int j, int k) {
    return ((size_t)(k + b->h[2]) * b->s[1] + (size_t)(j + b->h[1])) *
               b->s[0] + (size_t)(i + b->h[0]);
}
static inline int XIi(const orc_case* c, int l, int d) {
    return d < c->dim ? (int)c->XI[l * c->dim + d] : 0;
}

/* ---- CalcBGKFeq, T=1, polyOrder=2: model_host_device.h:34-57, 58-83 --- */
static inline double feq2d(const orc_case* c, int l, double rho, double u,
                           double v) {
    const double CS = c->CS;
    const double* XI = c->XI;
    const double T = 1;
    double cu = (CS * XI[l * 2] * u + CS * XI[l * 2 + 1] * v);
    double c2 = (CS * XI[l * 2] * CS * XI[l * 2] +
                 CS * XI[l * 2 + 1] * CS * XI[l * 2 + 1]);
    double cu2 = cu * cu;
    double u2 = u * u + v * v;
    double res = 1.0 + cu + 0.5 * (cu2 - u2 + (T - 1.0) * (c2 - 2));
    return c->W[l] * rho * res;
}
static inline double feq3d(const orc_case* c, int l, double rho, double u,
                           double v, double w) {
    const double CS = c->CS;
    const double* XI = c->XI;
    const double T = 1;
    double cu = (CS * XI[l * 3] * u + CS * XI[l * 3 + 1] * v +
                 CS * XI[l * 3 + 2] * w);
    double c2 = (CS * XI[l * 3] * CS * XI[l * 3] +
                 CS * XI[l * 3 + 1] * CS * XI[l * 3 + 1] +
                 CS * XI[l * 3 + 2] * CS * XI[l * 3 + 2]);
    double cu2 = cu * cu;
    double u2 = u * u + v * v + w * w;
    double res = 1.0 + cu + 0.5 * (cu2 - u2 + (T - 1.0) * (c2 - 3));
    return c->W[l] * rho * res;
}
static inline double feq(const orc_case* c, int l, double rho, double u,
                         double v, double w) {
    return c->dim == 2 ? feq2d(c, l, rho, u, v) : feq3d(c, l, rho, u, v, w);
}

/* ---- CalcBGKFeq with temperature and polyOrder, model_host_device.h:34-57
 * (2D) and 58-83 (3D), every term and its operation order; LATTDIM is an
 * int in the reference, so c2 - LATTDIM etc. convert it to double ------- */
static inline double feq_t2d(const orc_case* c, int l, double rho, double u,
                             double v, double T, int polyOrder) {
    const double CS = c->CS;
    const double* XI = c->XI;
    const int LATTDIM = 2;
    double cu = (CS * XI[l * LATTDIM] * u + CS * XI[l * LATTDIM + 1] * v);
    double c2 = (CS * XI[l * LATTDIM] * CS * XI[l * LATTDIM] +
                 CS * XI[l * LATTDIM + 1] * CS * XI[l * LATTDIM + 1]);
    double cu2 = cu * cu;
    double u2 = u * u + v * v;
    double res = 1.0 + cu + 0.5 * (cu2 - u2 + (T - 1.0) * (c2 - LATTDIM));
    if (polyOrder >= 3)
        res = res + cu * (cu2 - 3.0 * u2 + 3.0 * (T - 1.0) * (c2 - LATTDIM - 2.0)) / 6.0;
    if (polyOrder >= 4)
        res = res + (cu2 * cu2 - 6.0 * cu2 * u2 + 3.0 * u2 * u2) / 24.0 +
              (T - 1.0) * ((c2 - (LATTDIM + 2)) * (cu2 - u2) - 2.0 * cu2) / 4.0 +
              (T - 1.0) * (T - 1.0) *
                  (c2 * c2 - 2.0 * (LATTDIM + 2) * c2 + LATTDIM * (LATTDIM + 2)) / 8.0;
    return c->W[l] * rho * res;
}
static inline double feq_t3d(const orc_case* c, int l, double rho, double u,
                             double v, double w, double T, int polyOrder) {
    const double CS = c->CS;
    const double* XI = c->XI;
    const int LATTDIM = 3;
    double cu = (CS * XI[l * LATTDIM] * u + CS * XI[l * LATTDIM + 1] * v +
                 CS * XI[l * LATTDIM + 2] * w);
    double c2 = (CS * XI[l * LATTDIM] * CS * XI[l * LATTDIM] +
                 CS * XI[l * LATTDIM + 1] * CS * XI[l * LATTDIM + 1] +
                 CS * XI[l * LATTDIM + 2] * CS * XI[l * LATTDIM + 2]);
    double cu2 = cu * cu;
    double u2 = u * u + v * v + w * w;
    double res = 1.0 + cu + 0.5 * (cu2 - u2 + (T - 1.0) * (c2 - LATTDIM));
    if (polyOrder >= 3)
        res = res + cu * (cu2 - 3.0 * u2 + 3.0 * (T - 1.0) * (c2 - LATTDIM - 2.0)) / 6.0;
    if (polyOrder >= 4)
        res = res + (cu2 * cu2 - 6.0 * cu2 * u2 + 3.0 * u2 * u2) / 24.0 +
              (T - 1.0) * ((c2 - (LATTDIM + 2)) * (cu2 - u2) - 2.0 * cu2) / 4.0 +
              (T - 1.0) * (T - 1.0) *
                  (c2 * c2 - 2.0 * (LATTDIM + 2) * c2 + LATTDIM * (LATTDIM + 2)) / 8.0;
    return c->W[l] * rho * res;
}
double orc_feq_t(orc_case* c, int l, double rho, double u, double v, double w,
                 double T, int poly_order) {
    return c->dim == 2 ? feq_t2d(c, l, rho, u, v, T, poly_order)
                       : feq_t3d(c, l, rho, u, v, w, T, poly_order);
}

/* ---- creation -------------------------------------------------------- */
orc_case* orc_create(int space_dim, int lattice) {
    if (!((space_dim == 2 && lattice == 0) ||
          (space_dim == 3 && (lattice == 1 || lattice == 2))))
        return NULL;
    orc_case* c = (orc_case*)calloc(1, sizeof(orc_case));
    c->dim = space_dim;
    setup_lattice(c, lattice);
    c->tau = 1.0;
    c->dt = 1.0; /* DT{1} flowfield.cpp:88 */
    c->force_type = 0;
    c->scheme = SCHEME_STREAM_COLLISION;
    return c;
}

void orc_destroy(orc_case* c) {
    if (!c) return;
    for (int ib = 0; ib < c->nb; ib++) {
        OBlock* b = &c->b[ib];
        free(b->f); free(b->fs); free(b->nt); free(b->vg); free(b->solid); free(b->temp);
        for (int v = 0; v < 4; v++) { free(b->m[v]); free(b->mc[v]); }
    }
    free(c);
}

int orc_add_block(orc_case* c, int id, int nx, int ny, int nz) {
    if (c->nb >= MAXB) return fail(c, "too many blocks");
    if (id != c->nb) return fail(c, "blocks must be added in id order");
    if (nx < 1 || ny < 1 || (c->dim == 3 && nz < 1))
        return fail(c, "bad block size");
    OBlock* b = &c->b[c->nb++];
    memset(b, 0, sizeof *b);
    b->id = id;
    b->n[0] = nx; b->n[1] = ny; b->n[2] = c->dim == 3 ? nz : 1;
    b->h[0] = 1; b->h[1] = 1; b->h[2] = c->dim == 3 ? 1 : 0;
    for (int a = 0; a < 3; a++) b->s[a] = b->n[a] + 2 * b->h[a];
    b->nalloc = (size_t)b->s[0] * b->s[1] * b->s[2];
    /* ops_decl_dat without data: zero-filled storage */
    b->f = (double*)calloc(b->nalloc * c->Q, sizeof(double));
    b->fs = (double*)calloc(b->nalloc * c->Q, sizeof(double));
    b->nt = (int*)calloc(b->nalloc, sizeof(int));
    b->vg = (int*)calloc(b->nalloc, sizeof(int));
    for (int v = 0; v < 4; v++) {
        b->m[v] = (double*)calloc(b->nalloc, sizeof(double));
        b->mc[v] = (double*)calloc(b->nalloc, sizeof(double));
    }
    /* Variable_T: the thermal collision's temperature, 1 until set */
    b->temp = (double*)malloc(b->nalloc * sizeof(double));
    for (size_t n = 0; n < b->nalloc; n++) b->temp[n] = 1.0;
    return 0;
}

int orc_add_connection(orc_case* c, int from, int fsurf, int to, int tsurf,
                       int type) {
    if (c->nlink >= MAXLINK) return fail(c, "too many connections");
    if (from < 0 || from >= c->nb || to < 0 || to >= c->nb)
        return fail(c, "connection names an undefined block");
    for (int l = 0; l < c->nlink; l++) /* Block::AddNeighbor block.cpp:141 */
        if (c->link[l].from == from && c->link[l].fsurf == fsurf)
            return fail(c, "a neighbour is already defined on this surface");
    OLink L = {from, fsurf, to, tsurf, type, 0};
    c->link[c->nlink++] = L;
    return 0;
}

/* face surface -> axis and outward side; -1 for edge surfaces */
static int face_axis(int s, int* side) {
    *side = 0;
    switch (s) {
        case BS_LEFT: *side = -1; return 0;
        case BS_RIGHT: *side = 1; return 0;
        case BS_BOTTOM: *side = -1; return 1;
        case BS_TOP: *side = 1; return 1;
        case BS_BACK: *side = -1; return 2;
        case BS_FRONT: *side = 1; return 2;
        default: return -1;
    }
}

/* Extension with no reference counterpart (the reference links same-axis
 * faces only, field.h:207-214): fStage of the from-face plane goes into the
 * halo beyond the to-face, turned by the quarter rotation R about the third
 * axis that takes the from-face's outward normal n_f to the to-face's inward
 * normal -n_t.  Populations move with it: f'_{q'} = f_q where c_q' = R c_q. */
int orc_add_rotated_connection(orc_case* c, int from, int fsurf, int to,
                               int tsurf, int type) {
    int sf, st;
    const int af = face_axis(fsurf, &sf), at = face_axis(tsurf, &st);
    if (af < 0 || at < 0 || af == at || af >= c->dim || at >= c->dim)
        return fail(c, "a rotated connection joins two perpendicular faces");
    if (orc_add_connection(c, from, fsurf, to, tsurf, type)) return -1;
    c->link[c->nlink - 1].rot = 1;
    return 0;
}

/* R (new = R old) of a link: identity for reference links */
static void link_rotation(const OLink* L, int R[3][3]) {
    memset(R, 0, 9 * sizeof(int));
    if (!L->rot) {
        for (int k = 0; k < 3; k++) R[k][k] = 1;
        return;
    }
    int sf, st;
    const int af = face_axis(L->fsurf, &sf), at = face_axis(L->tsurf, &st);
    const int a3 = 3 - af - at;
    const int sigma = -sf * st;
    R[at][af] = sigma;  /* n_f = sf e_af  ->  -st e_at */
    R[af][at] = -sigma; /* det R = +1 */
    R[a3][a3] = 1;
}

/* target node of from-face-plane node p (transverse halo rows included) */
static void link_target(const orc_case* c, const OLink* L, const int p[3],
                        int out[3]) {
    const OBlock* t = &c->b[L->to];
    int R[3][3], sf, st;
    link_rotation(L, R);
    const int af = face_axis(L->fsurf, &sf);
    int at = face_axis(L->tsurf, &st);
    if (!L->rot) { at = af; st = -sf; } /* the reference ignores tsurf */
    for (int k = 0; k < 3; k++) {
        if (k == af) {
            out[at] = st > 0 ? t->n[at] : -1;
            continue;
        }
        for (int j = 0; j < 3; j++) {
            if (R[j][k] == 1) out[j] = p[k];
            if (R[j][k] == -1) out[j] = t->n[j] - 1 - p[k];
        }
    }
    if (c->dim == 2) out[2] = 0;
}

/* population map of a link: qmap[q] = q' with c_q' = R c_q */
static void link_qmap(const orc_case* c, const OLink* L, int* qmap) {
    int R[3][3];
    link_rotation(L, R);
    const int D = c->dim;
    for (int q = 0; q < c->Q; q++) {
        int v[3] = {0, 0, 0};
        for (int j = 0; j < D; j++)
            for (int k = 0; k < D; k++) v[j] += R[j][k] * (int)c->XI[q * D + k];
        qmap[q] = -1;
        for (int r = 0; r < c->Q; r++) {
            int same = 1;
            for (int j = 0; j < D; j++) same = same && (int)c->XI[r * D + j] == v[j];
            if (same) { qmap[q] = r; break; }
        }
    }
}

int orc_set_collision(orc_case* c, double tau_ref, int force_type,
                      int force_macros) {
    if (force_type != 0 && force_type != 1)
        return fail(c, "only BodyForce_None and BodyForce_1st are restated");
    c->tau = tau_ref;
    c->force_type = force_type;
    c->force_macros = force_macros;
    return 0;
}

int orc_set_collision_type(orc_case* c, int type) {
    if (type != COLL_ISOTHERMAL2ND && type != COLL_THERMAL4TH)
        return fail(c, "collision type %d is not restated", type);
    c->collision = type;
    return 0;
}

int orc_set_temperature(orc_case* c, int block, const double* T) {
    if (block < 0 || block >= c->nb) return fail(c, "undefined block");
    OBlock* b = &c->b[block];
    size_t p = 0;
    for (int k = 0; k < b->n[2]; k++)
        for (int j = 0; j < b->n[1]; j++)
            for (int i = 0; i < b->n[0]; i++) b->temp[nidx(b, i, j, k)] = T[p++];
    return 0;
}

int orc_add_boundary(orc_case* c, int block, int surface, int scheme,
                     const double* vars, int nvars, int vtype) {
    if (c->nbc >= MAXBC) return fail(c, "too many boundaries");
    if (block < 0 || block >= c->nb) return fail(c, "undefined block");
    if (nvars > 8) return fail(c, "too many given variables");
    OBC* bc = &c->bc[c->nbc++];
    memset(bc, 0, sizeof *bc);
    bc->block = block; bc->surface = surface; bc->scheme = scheme;
    bc->vtype = vtype; bc->nvars = nvars;
    for (int i = 0; i < nvars; i++) bc->vars[i] = vars[i];
    return 0;
}

/* ---- Block::RangeStart/RangeEnd, block.cpp:14-97 ---------------------- */
static void surface_range(const orc_case* c, const OBlock* b, int s,
                          int r[6]) {
    for (int a = 0; a < 3; a++) { r[2 * a] = 0; r[2 * a + 1] = b->n[a]; }
    if (s == BS_RIGHT || s == BS_RIGHTTOP || s == BS_RIGHTBOTTOM ||
        (c->dim == 3 && (s == BS_RIGHTFRONT || s == BS_RIGHTBACK)))
        r[0] = b->n[0] - 1;
    if (s == BS_LEFT || s == BS_LEFTBOTTOM || s == BS_LEFTTOP ||
        (c->dim == 3 && (s == BS_LEFTFRONT || s == BS_LEFTBACK)))
        r[1] = 1;
    /* block.cpp:35 tests "OP_3D", never defined: TopBack/TopFront keep the
     * full y range (parity quirk, reproduced). */
    if (s == BS_TOP || s == BS_LEFTTOP || s == BS_RIGHTTOP)
        r[2] = b->n[1] - 1;
    if (s == BS_BOTTOM || s == BS_LEFTBOTTOM || s == BS_RIGHTBOTTOM ||
        (c->dim == 3 && (s == BS_BOTTOMBACK || s == BS_BOTTOMFRONT)))
        r[3] = 1;
    if (c->dim == 3) {
        if (s == BS_FRONT || s == BS_LEFTFRONT || s == BS_RIGHTFRONT ||
            s == BS_TOPFRONT || s == BS_BOTTOMFRONT)
            r[4] = b->n[2] - 1;
        if (s == BS_BACK || s == BS_LEFTBACK || s == BS_RIGHTBACK ||
            s == BS_TOPBACK || s == BS_BOTTOMBACK)
            r[5] = 1;
    }
}

static void set_int_range(OBlock* b, int* field, const int r[6], int val) {
    for (int k = r[4]; k < r[5]; k++)
        for (int j = r[2]; j < r[3]; j++)
            for (int i = r[0]; i < r[1]; i++) field[nidx(b, i, j, k)] = val;
}

/* the six halo slabs around a block (flowfield_wrapper.cpp:303-390, the
 * same ranges again in SetBulkandHaloNodesType :671-788) */
static void set_halo_slabs(const orc_case* c, OBlock* b, int* field,
                           int val) {
    const int nx = b->n[0], ny = b->n[1], nz = b->n[2];
    const int zl = c->dim == 3 ? -1 : 0, zh = c->dim == 3 ? nz + 1 : 1;
    int r[6];
    int rb[6] = {-1, nx + 1, -1, 0, zl, zh};      set_int_range(b, field, rb, val);
    int rt[6] = {-1, nx + 1, ny, ny + 1, zl, zh}; set_int_range(b, field, rt, val);
    int rl[6] = {-1, 0, -1, ny + 1, zl, zh};      set_int_range(b, field, rl, val);
    int rr[6] = {nx, nx + 1, -1, ny + 1, zl, zh}; set_int_range(b, field, rr, val);
    if (c->dim == 3) {
        int rk[6] = {-1, nx + 1, -1, ny + 1, -1, 0};    set_int_range(b, field, rk, val);
        int rf[6] = {-1, nx + 1, -1, ny + 1, nz, nz + 1}; set_int_range(b, field, rf, val);
    }
    (void)r;
}

/* SetBlockGeometryProperty, flowfield_wrapper.cpp:291-623 */
static void set_geometry(orc_case* c, OBlock* b) {
    const int nx = b->n[0], ny = b->n[1], nz = b->n[2];
    int r[6];
    int bulk[6] = {1, nx - 1, 1, ny - 1, c->dim == 3 ? 1 : 0,
                   c->dim == 3 ? nz - 1 : 1};
    set_int_range(b, b->vg, bulk, VG_FLUID);
    set_halo_slabs(c, b, b->vg, VG_IMMERSEDSOLID);
    surface_range(c, b, BS_BOTTOM, r); set_int_range(b, b->vg, r, VG_JP);
    surface_range(c, b, BS_TOP, r);    set_int_range(b, b->vg, r, VG_JM);
    surface_range(c, b, BS_LEFT, r);   set_int_range(b, b->vg, r, VG_IP);
    surface_range(c, b, BS_RIGHT, r);  set_int_range(b, b->vg, r, VG_IM);
    if (c->dim == 3) {
        surface_range(c, b, BS_BACK, r);  set_int_range(b, b->vg, r, VG_KP);
        surface_range(c, b, BS_FRONT, r); set_int_range(b, b->vg, r, VG_KM);
    }
    if (c->dim == 2) {
        int a[6] = {0, 1, 0, 1, 0, 1};           set_int_range(b, b->vg, a, 10200);
        int d[6] = {0, 1, ny - 1, ny, 0, 1};     set_int_range(b, b->vg, d, 10210);
        int e[6] = {nx - 1, nx, ny - 1, ny, 0, 1}; set_int_range(b, b->vg, e, 11210);
        int g[6] = {nx - 1, nx, 0, 1, 0, 1};     set_int_range(b, b->vg, g, 11200);
        return;
    }
    /* 12 edges */
    const int E[12][7] = {
        {0, 1, 0, 1, 0, nz, 10200},           {0, 1, ny - 1, ny, 0, nz, 10210},
        {nx - 1, nx, ny - 1, ny, 0, nz, 11210}, {nx - 1, nx, 0, 1, 0, nz, 11200},
        {0, 1, 0, ny, 0, 1, 10300},           {0, 1, 0, ny, nz - 1, nz, 10310},
        {nx - 1, nx, 0, ny, nz - 1, nz, 11310}, {nx - 1, nx, 0, ny, 0, 1, 11300},
        {0, nx, 0, 1, 0, 1, 20300},           {0, nx, 0, 1, nz - 1, nz, 20310},
        {0, nx, ny - 1, ny, nz - 1, nz, 21310}, {0, nx, ny - 1, ny, 0, 1, 21300}};
    for (int e = 0; e < 12; e++) set_int_range(b, b->vg, E[e], E[e][6]);
    /* 8 corners */
    const int C8[8][7] = {
        {0, 1, 0, 1, 0, 1, 1020300},
        {0, 1, 0, 1, nz - 1, nz, 1020310},
        {0, 1, ny - 1, ny, 0, 1, 1021300},
        {0, 1, ny - 1, ny, nz - 1, nz, 1021310},
        {nx - 1, nx, 0, 1, 0, 1, 1120300},
        {nx - 1, nx, 0, 1, nz - 1, nz, 1120310},
        {nx - 1, nx, ny - 1, ny, 0, 1, 1121300},
        {nx - 1, nx, ny - 1, ny, nz - 1, nz, 1121310}};
    for (int e = 0; e < 8; e++) set_int_range(b, b->vg, C8[e], C8[e][6]);
}

/* SetBulkandHaloNodesType, flowfield_wrapper.cpp:671-788 */
static void set_bulk_halo_types(orc_case* c, OBlock* b) {
    int bulk[6] = {1, b->n[0] - 1, 1, b->n[1] - 1, c->dim == 3 ? 1 : 0,
                   c->dim == 3 ? b->n[2] - 1 : 1};
    set_int_range(b, b->nt, bulk, VT_FLUID);
    set_halo_slabs(c, b, b->nt, VT_IMMERSEDSOLID);
}

/* SetBoundaryNodeType, flowfield_wrapper.cpp:625-669 */
static void set_boundary_types(orc_case* c) {
    for (int ib = 0; ib < c->nbc; ib++) {
        OBC* bc = &c->bc[ib];
        OBlock* b = &c->b[bc->block];
        int r[6];
        surface_range(c, b, bc->surface, r);
        if (bc->vtype == VT_VIRTUALBOUNDARY) {
            if (bc->surface == BS_LEFT || bc->surface == BS_RIGHT) {
                r[2]++; r[3]--;
                if (c->dim == 3) { r[4]++; r[5]--; }
            }
            if (bc->surface == BS_TOP || bc->surface == BS_BOTTOM) {
                r[0]++; r[1]--;
                if (c->dim == 3) { r[4]++; r[5]--; }
            }
            if (c->dim == 3 &&
                (bc->surface == BS_FRONT || bc->surface == BS_BACK)) {
                r[0]++; r[1]--; r[2]++; r[3]--;
            }
        }
        set_int_range(b, b->nt, r, bc->vtype);
    }
}

static void copy_current_macros(orc_case* c) {
    for (int ib = 0; ib < c->nb; ib++) {
        OBlock* b = &c->b[ib];
        for (int v = 0; v < 4; v++)
            memcpy(b->mc[v], b->m[v], b->nalloc * sizeof(double));
    }
}

/* Extension with no reference counterpart: in-domain obstacles.  Obstacle
 * nodes are ImmersedSolid (no macros, collision or streaming, like the
 * reference's solid nodes, model_kernel.inc:420, scheme_kernel.inc:1418);
 * a node that pulls every population (Fluid, MDPeriodic, VirtualBoundary)
 * and whose pull source x - c_q is an obstacle node takes the half-way
 * bounce-back value f_q(x) = fs_opp(q)(x) of a resting wall.  Sources in a
 * halo are resolved through the block connections.  orc_prepare refuses
 * obstacles on nodes that are not Fluid or MDPeriodic, or next to nodes
 * that do not pull every population. */
int orc_set_obstacles(orc_case* c, int block, const unsigned char* solid) {
    if (c->prepared) return fail(c, "obstacles are set before orc_prepare");
    if (block < 0 || block >= c->nb) return fail(c, "undefined block");
    OBlock* b = &c->b[block];
    if (!b->solid) b->solid = (unsigned char*)calloc(b->nalloc, 1);
    size_t p = 0;
    for (int k = 0; k < b->n[2]; k++)
        for (int j = 0; j < b->n[1]; j++)
            for (int i = 0; i < b->n[0]; i++, p++)
                b->solid[nidx(b, i, j, k)] = solid[p] ? 1 : 0;
    return 0;
}

/* obstacle flags of halo nodes, copied through the connections like fStage */
static void obstacle_halos(orc_case* c) {
    for (int l = 0; l < c->nlink; l++) {
        const OLink* L = &c->link[l];
        OBlock* a = &c->b[L->from];
        OBlock* t = &c->b[L->to];
        if (!a->solid) continue;
        if (!t->solid) t->solid = (unsigned char*)calloc(t->nalloc, 1);
        int side;
        const int ax = face_axis(L->fsurf, &side);
        int lo[3] = {0, 0, 0}, hi[3] = {a->n[0], a->n[1], a->n[2]};
        lo[ax] = side > 0 ? a->n[ax] - 1 : 0;
        hi[ax] = lo[ax] + 1;
        for (int k = lo[2]; k < hi[2]; k++)
            for (int j = lo[1]; j < hi[1]; j++)
                for (int i = lo[0]; i < hi[0]; i++) {
                    const int p[3] = {i, j, k};
                    int d[3];
                    link_target(c, L, p, d);
                    t->solid[nidx(t, d[0], d[1], d[2])] = a->solid[nidx(a, i, j, k)];
                }
    }
}

static int full_pull(int vt) {
    return vt == VT_FLUID || vt == VT_MDPERIODIC || vt == VT_VIRTUALBOUNDARY;
}

static int mark_obstacles(orc_case* c) {
    obstacle_halos(c);
    for (int ib = 0; ib < c->nb; ib++) {
        OBlock* b = &c->b[ib];
        if (!b->solid) continue;
        for (int k = 0; k < b->n[2]; k++)
            for (int j = 0; j < b->n[1]; j++)
                for (int i = 0; i < b->n[0]; i++) {
                    const size_t n = nidx(b, i, j, k);
                    if (b->solid[n]) {
                        if (b->nt[n] != VT_FLUID && b->nt[n] != VT_MDPERIODIC)
                            return fail(c, "obstacle on a node that is not Fluid or MDPeriodic");
                        continue;
                    }
                    /* a node that pulls from an obstacle must pull everything */
                    for (int q = 0; q < c->Q; q++) {
                        const size_t m = nidx(b, i - XIi(c, q, 0), j - XIi(c, q, 1),
                                              k - XIi(c, q, 2));
                        if (b->solid[m] && !full_pull(b->nt[n]))
                            return fail(c, "obstacle next to a boundary node");
                    }
                }
    }
    for (int ib = 0; ib < c->nb; ib++) {
        OBlock* b = &c->b[ib];
        if (!b->solid) continue;
        for (int k = 0; k < b->n[2]; k++)
            for (int j = 0; j < b->n[1]; j++)
                for (int i = 0; i < b->n[0]; i++)
                    if (b->solid[nidx(b, i, j, k)]) b->nt[nidx(b, i, j, k)] = VT_IMMERSEDSOLID;
    }
    return 0;
}

int orc_prepare(orc_case* c) {
    for (int l = 0; l < c->nlink; l++) {
        const OLink* L = &c->link[l];
        const OBlock* a = &c->b[L->from];
        const OBlock* t = &c->b[L->to];
        int side;
        const int ax = face_axis(L->fsurf, &side);
        if (ax < 0 || ax >= c->dim)
            return fail(c, "a connection joins face surfaces");
        int R[3][3];
        link_rotation(L, R);
        for (int d = 0; d < c->dim; d++)
            for (int j = 0; j < c->dim; j++)
                if (d != ax && R[j][d] != 0 && a->n[d] != t->n[j])
                    return fail(c, "connected blocks differ in a transverse size");
    }
    for (int ib = 0; ib < c->nb; ib++) {
        set_geometry(c, &c->b[ib]);
        set_bulk_halo_types(c, &c->b[ib]);
    }
    set_boundary_types(c);
    if (mark_obstacles(c)) return -1;
    copy_current_macros(c);
    c->prepared = 1;
    return 0;
}

int orc_set_macros(orc_case* c, int block, const double* rho,
                   const double* u, const double* v, const double* w) {
    if (block < 0 || block >= c->nb) return fail(c, "undefined block");
    OBlock* b = &c->b[block];
    const double* src[4] = {rho, u, v, w};
    size_t p = 0;
    for (int k = 0; k < b->n[2]; k++)
        for (int j = 0; j < b->n[1]; j++)
            for (int i = 0; i < b->n[0]; i++, p++) {
                size_t n = nidx(b, i, j, k);
                for (int var = 0; var < 1 + c->dim; var++)
                    b->m[var][n] = src[var][p];
            }
    return 0;
}

void orc_set_dt(orc_case* c, double dt) { c->dt = dt; }

int orc_set_accel(orc_case* c, int block, const double* g) {
    if (block < 0 || block >= c->nb) return fail(c, "undefined block");
    for (int d = 0; d < 3; d++) c->b[block].g[d] = d < c->dim ? g[d] : 0.0;
    return 0;
}

/* KerInitialiseBGK2nd[3D], model_kernel.inc:193-219, 352-380, over the
 * WholeRange, then CopyCurrentMacroVar (model_wrapper.cpp:387-389) */
int orc_init_feq(orc_case* c) {
    if (!c->prepared) return fail(c, "call orc_prepare first");
    const int Q = c->Q;
    for (int ib = 0; ib < c->nb; ib++) {
        OBlock* b = &c->b[ib];
        for (int k = 0; k < b->n[2]; k++)
            for (int j = 0; j < b->n[1]; j++)
                for (int i = 0; i < b->n[0]; i++) {
                    size_t n = nidx(b, i, j, k);
                    if (b->nt[n] == VT_IMMERSEDSOLID) continue;
                    for (int q = 0; q < Q; q++)
                        b->f[n * Q + q] = feq(c, q, b->m[0][n], b->m[1][n],
                                              b->m[2][n], b->m[3][n]);
                }
    }
    copy_current_macros(c);
    return 0;
}

/* ---- the step ---------------------------------------------------------- */

/* KerCalcDensity[3D], KerCalcU/V/W[3D] and the *Force variants,
 * model_kernel.inc:57-186, 570-759, driven by model_wrapper.cpp:112-275 */
void orc_update_macros(orc_case* c) {
    const int Q = c->Q, D = c->dim;
    const double CS = c->CS;
    const double* XI = c->XI;
    for (int ib = 0; ib < c->nb; ib++) {
        OBlock* b = &c->b[ib];
        const int nk = b->n[2];
        OMP_FOR
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < b->n[1]; j++)
                for (int i = 0; i < b->n[0]; i++) {
                    size_t n = nidx(b, i, j, k);
                    int vt = b->nt[n];
                    if (vt == VT_IMMERSEDSOLID) continue;
                    const double* f = &b->f[n * Q];
                    double rho = 0;
                    for (int q = 0; q < Q; q++) rho += f[q];
                    b->m[0][n] = rho;
                    for (int d = 0; d < D; d++) {
                        double u = 0;
                        for (int q = 0; q < Q; q++)
                            u += CS * XI[q * D + d] * f[q];
                        u /= b->m[0][n];
                        if (c->force_macros &&
                            (vt == VT_FLUID || vt == VT_MDPERIODIC))
                            u += (c->dt * b->g[d] / 2);
                        b->m[1 + d][n] = u;
                    }
                }
    }
}

/* CopyBlockEnvelopDistribution, flowfield_wrapper.cpp:113-232 */
static void envelope_copy(orc_case* c) {
    const int Q = c->Q;
    const int faces[6] = {BS_LEFT, BS_RIGHT, BS_BOTTOM, BS_TOP, BS_BACK,
                          BS_FRONT};
    for (int ib = 0; ib < c->nb; ib++) {
        OBlock* b = &c->b[ib];
        for (int fidx = 0; fidx < (c->dim == 3 ? 6 : 4); fidx++) {
            int r[6];
            surface_range(c, b, faces[fidx], r);
            for (int k = r[4]; k < r[5]; k++)
                for (int j = r[2]; j < r[3]; j++)
                    for (int i = r[0]; i < r[1]; i++) {
                        size_t n = nidx(b, i, j, k);
                        for (int q = 0; q < Q; q++)
                            b->fs[n * Q + q] = b->f[n * Q + q];
                    }
        }
    }
}

/* KerCalcBodyForce1ST[3D] / KerCalcBodyForceNone[3D],
 * model_kernel.inc:312-348, 504-541; CalcBodyForce model_host_device.h:25-32 */
static void body_force(orc_case* c) {
    const int Q = c->Q, D = c->dim;
    const double CS = c->CS;
    const double* XI = c->XI;
    for (int ib = 0; ib < c->nb; ib++) {
        OBlock* b = &c->b[ib];
        OMP_FOR
        for (int k = 0; k < b->n[2]; k++)
            for (int j = 0; j < b->n[1]; j++)
                for (int i = 0; i < b->n[0]; i++) {
                    size_t n = nidx(b, i, j, k);
                    int vt = b->nt[n];
                    if (!(vt == VT_FLUID || vt == VT_MDPERIODIC)) continue;
                    for (int q = 0; q < Q; q++) {
                        if (c->force_type == 1) {
                            double cf = 0;
                            for (int d = 0; d < D; d++)
                                cf += CS * XI[q * D + d] * b->g[d];
                            b->fs[n * Q + q] = c->W[q] * b->m[0][n] * cf;
                        } else {
                            b->fs[n * Q + q] = 0;
                        }
                    }
                }
    }
}

/* KerCollideBGKThermal[3D], model_kernel.inc:267-310, 457-502: per node
 * tau = tauRef / (rho sqrt(T)), the 4th-order equilibrium at temperature T,
 * and the update f - w (f - feq) (+ tau w F at Fluid and MDPeriodic) */
static void collide_thermal(orc_case* c) {
    const int Q = c->Q;
    const double dt = c->dt;
    for (int ib = 0; ib < c->nb; ib++) {
        OBlock* b = &c->b[ib];
        OMP_FOR
        for (int k = 0; k < b->n[2]; k++)
            for (int j = 0; j < b->n[1]; j++)
                for (int i = 0; i < b->n[0]; i++) {
                    size_t n = nidx(b, i, j, k);
                    int vt = b->nt[n];
                    if (vt == VT_IMMERSEDSOLID) continue;
                    const double rho = b->m[0][n], u = b->m[1][n],
                                 v = b->m[2][n], w = b->m[3][n];
                    const double T = b->temp[n];
                    const int polyOrder = 4;
                    const double tau = c->tau / (rho * sqrt(T));
                    const double dtOvertauPlusdt = dt / (tau + 0.5 * dt);
                    double* fs = &b->fs[n * Q];
                    const double* f = &b->f[n * Q];
                    for (int q = 0; q < Q; q++) {
                        const double fe = c->dim == 2
                                              ? feq_t2d(c, q, rho, u, v, T, polyOrder)
                                              : feq_t3d(c, q, rho, u, v, w, T, polyOrder);
                        if (vt == VT_FLUID || vt == VT_MDPERIODIC)
                            fs[q] = f[q] - dtOvertauPlusdt * (f[q] - fe) +
                                    tau * dtOvertauPlusdt * fs[q];
                        else
                            fs[q] = f[q] - dtOvertauPlusdt * (f[q] - fe);
                    }
                }
    }
}

/* KerCollideBGKIsothermal[3D], model_kernel.inc:221-265, 410-455 */
static void collide(orc_case* c) {
    if (c->collision == COLL_THERMAL4TH) {
        collide_thermal(c);
        return;
    }
    const int Q = c->Q;
    const double tau = c->tau, dt = c->dt;
    const double dtOvertauPlusdt = dt / (tau + 0.5 * dt);
    for (int ib = 0; ib < c->nb; ib++) {
        OBlock* b = &c->b[ib];
        OMP_FOR
        for (int k = 0; k < b->n[2]; k++)
            for (int j = 0; j < b->n[1]; j++)
                for (int i = 0; i < b->n[0]; i++) {
                    size_t n = nidx(b, i, j, k);
                    int vt = b->nt[n];
                    if (vt == VT_IMMERSEDSOLID) continue;
                    const double rho = b->m[0][n], u = b->m[1][n],
                                 v = b->m[2][n], w = b->m[3][n];
                    double* fs = &b->fs[n * Q];
                    const double* f = &b->f[n * Q];
                    for (int q = 0; q < Q; q++) {
                        const double fe = feq(c, q, rho, u, v, w);
                        if (vt == VT_FLUID || vt == VT_MDPERIODIC)
                            fs[q] = fe + (1 - dtOvertauPlusdt) * (f[q] - fe) +
                                    tau * dtOvertauPlusdt * fs[q];
                        else
                            fs[q] = fe + (1 - dtOvertauPlusdt) * (f[q] - fe);
                    }
                }
    }
}

/* ops_halo_transfer of fStage over the halos declared by Field::CreateHalos,
 * field.h:215-395: blocks in id order, neighbours in BoundarySurface order,
 * slabs one node deep that span the transverse halos. */
static void halo_transfer_field(orc_case* c, int use_f);
static void halo_transfer(orc_case* c) { halo_transfer_field(c, 0); }

/* use_f: the halos of f (the swap scheme registers f, scheme.cpp:102-105),
 * else those of fStage (scheme.cpp:96-99) */
static void halo_transfer_field(orc_case* c, int use_f) {
    const int Q = c->Q;
    for (int ib = 0; ib < c->nb; ib++) {
        const int order[6] = {BS_LEFT, BS_RIGHT, BS_TOP, BS_BOTTOM, BS_FRONT,
                              BS_BACK};
        for (int oi = 0; oi < 6; oi++) {
            for (int l = 0; l < c->nlink; l++) {
                const OLink* L = &c->link[l];
                if (L->from != ib || L->fsurf != order[oi]) continue;
                OBlock* a = &c->b[L->from];
                OBlock* t = &c->b[L->to];
                const int hz = c->dim == 3 ? 1 : 0;
                int sz[3] = {a->n[0] + 2, a->n[1] + 2, a->n[2] + 2 * hz};
                int src[3] = {-1, -1, -hz};
                int side;
                const int ax = face_axis(L->fsurf, &side);
                if (ax < 0) continue; /* rejected by orc_prepare */
                sz[ax] = 1;
                src[ax] = side > 0 ? a->n[ax] - 1 : 0;
                int qmap[19];
                link_qmap(c, L, qmap);
                for (int z = 0; z < sz[2]; z++)
                    for (int y = 0; y < sz[1]; y++)
                        for (int x = 0; x < sz[0]; x++) {
                            const int p[3] = {src[0] + x, src[1] + y, src[2] + z};
                            int d[3];
                            link_target(c, L, p, d);
                            const size_t ns = nidx(a, p[0], p[1], p[2]);
                            const size_t nd = nidx(t, d[0], d[1], d[2]);
                            double* tf = use_f ? t->f : t->fs;
                            const double* af = use_f ? a->f : a->fs;
                            for (int q = 0; q < Q; q++)
                                tf[nd * Q + qmap[q]] = af[ns * Q + q];
                        }
            }
        }
    }
}

/* ---- geometry decoding ------------------------------------------------- */
/* A geometry code concatenates two-digit face codes (10=IP, 11=IM, 20=JP,
 * 21=JM, 30=KP, 31=KM) and, for edges and corners, a trailing digit 0 (_I)
 * or 1 (_O); flowfield_host_device.h:77-144. */
static int vg_decode(int vg, int ax[3], int sg[3], int* inner) {
    int nf, codes[3];
    if (vg >= 10 && vg < 100) {
        nf = 1; codes[0] = vg; *inner = 1;
    } else if (vg >= 10000 && vg < 100000) {
        nf = 2; *inner = (vg % 10) == 0; vg /= 10;
        codes[0] = vg / 100; codes[1] = vg % 100;
    } else if (vg >= 1000000 && vg < 10000000) {
        nf = 3; *inner = (vg % 10) == 0; vg /= 10;
        codes[0] = vg / 10000; codes[1] = (vg / 100) % 100; codes[2] = vg % 100;
    } else {
        return 0;
    }
    for (int i = 0; i < nf; i++) {
        ax[i] = codes[i] / 10 - 1;
        sg[i] = (codes[i] % 10) == 0 ? 1 : -1;
        if (ax[i] < 0 || ax[i] > 2 || (codes[i] % 10) > 1) return 0;
    }
    return nf;
}

/* pull predicate of KerStream[3D] at boundary nodes,
 * scheme_kernel.inc:80-217 (2D), 1428-2049 (3D): every normal axis must
 * satisfy sign*c <= 0; all axes (_I) or any axis (_O). */
static int stream_pulls(int vg, const int cv[3]) {
    int ax[3], sg[3], inner;
    int nf = vg_decode(vg, ax, sg, &inner);
    if (nf == 0) return 0;
    int all = 1, any = 0;
    for (int i = 0; i < nf; i++) {
        int ok = sg[i] * cv[ax[i]] <= 0;
        all = all && ok;
        any = any || ok;
    }
    return (nf == 1 || inner) ? all : any;
}

/* FindBdyDvType[3D], boundary_host_device.h:47-183, 191-706, in components
 * along the inward normals cn[k] = sign_k * c_{axis_k}. */
static int dv_type(int vg, const int cv[3]) {
    int ax[3], sg[3], inner;
    int nf = vg_decode(vg, ax, sg, &inner);
    if (nf == 0) return DV_INVALID;
    int cn[3];
    for (int i = 0; i < nf; i++) cn[i] = sg[i] * cv[ax[i]];
    if (nf == 1) {
        if (cn[0] < 0) return DV_INCOMING;
        if (cn[0] == 0) return DV_PARALLEL;
        return DV_OUTGOING;
    }
    if (nf == 2) {
        if ((cn[0] <= 0 && cn[1] < 0) || (cn[0] < 0 && cn[1] == 0))
            return DV_INCOMING;
        if ((cn[0] >= 0 && cn[1] > 0) || (cn[0] > 0 && cn[1] == 0))
            return DV_OUTGOING;
        return DV_PARALLEL;
    }
    /* 3D corners: the reference names the wall-pointing set "Outgoing" */
    if (cv[0] == 0 && cv[1] == 0 && cv[2] == 0) return DV_PARALLEL;
    if (cn[0] <= 0 && cn[1] <= 0 && cn[2] <= 0) return DV_OUTGOING;
    if (cn[0] >= 0 && cn[1] >= 0 && cn[2] >= 0) return DV_INCOMING;
    return DV_PARALLEL;
}

/* KerStream / KerStream3D, scheme_kernel.inc:53-221, 1400-2052 */
static void stream(orc_case* c) {
    const int Q = c->Q;
    for (int ib = 0; ib < c->nb; ib++) {
        OBlock* b = &c->b[ib];
        OMP_FOR
        for (int k = 0; k < b->n[2]; k++)
            for (int j = 0; j < b->n[1]; j++)
                for (int i = 0; i < b->n[0]; i++) {
                    size_t n = nidx(b, i, j, k);
                    const int vt = b->nt[n], vg = b->vg[n];
                    for (int q = 0; q < Q; q++) {
                        int cv[3] = {XIi(c, q, 0), XIi(c, q, 1), XIi(c, q, 2)};
                        size_t src = nidx(b, i - cv[0], j - cv[1], k - cv[2]);
                        if (vt == VT_FLUID || vt == VT_MDPERIODIC ||
                            vt == VT_VIRTUALBOUNDARY) {
                            if (b->solid && b->solid[src])
                                b->f[n * Q + q] = b->fs[n * Q + c->OPP[q]];
                            else
                                b->f[n * Q + q] = b->fs[src * Q + q];
                            continue;
                        }
                        if (vt == VT_IMMERSEDSOLID) continue;
                        if (cv[0] == 0 && cv[1] == 0 && cv[2] == 0) {
                            b->f[n * Q + q] = b->fs[n * Q + q];
                            continue;
                        }
                        if (stream_pulls(vg, cv))
                            b->f[n * Q + q] = b->fs[src * Q + q];
                    }
                }
    }
}

/* KerCutCellEQMDiffuseRefl[3D], boundary_kernel.inc:534-613, 1184-1276 */
static void bc_eqm(orc_case* c, OBlock* b, const int r[6], const double* gv) {
    const int Q = c->Q, D = c->dim;
    const double CS = c->CS;
    const double* XI = c->XI;
    const double u = gv[0], v = gv[1], w = D == 3 ? gv[2] : 0.0;
    for (int k = r[4]; k < r[5]; k++)
        for (int j = r[2]; j < r[3]; j++)
            for (int i = r[0]; i < r[1]; i++) {
                size_t n = nidx(b, i, j, k);
                double* f = &b->f[n * Q];
                const int vg = b->vg[n];
                int outgoing[MAXQ], parallel[MAXQ];
                int nout = 0, npar = 0;
                double rhoIncoming = 0, rhoParallel = 0, deltaRho = 0;
                for (int q = 0; q < Q; q++) {
                    int cv[3] = {XIi(c, q, 0), XIi(c, q, 1), XIi(c, q, 2)};
                    const double cx = CS * XI[q * D];
                    const double cy = CS * XI[q * D + 1];
                    const double cz = D == 3 ? CS * XI[q * D + 2] : 0.0;
                    switch (dv_type(vg, cv)) {
                        case DV_INCOMING:
                            rhoIncoming += f[q];
                            break;
                        case DV_OUTGOING:
                            outgoing[nout++] = q;
                            if (D == 2)
                                deltaRho += (2 * c->W[q]) * (cx * u + cy * v);
                            else
                                deltaRho += (2 * c->W[q]) *
                                            (cx * u + cy * v + cz * w);
                            break;
                        case DV_PARALLEL:
                            parallel[npar++] = q;
                            rhoParallel += feq(c, q, 1, u, v, w);
                            break;
                        default:
                            break;
                    }
                }
                const double rhoWall =
                    2 * rhoIncoming / (1 - deltaRho - rhoParallel);
                for (int p = 0; p < npar; p++)
                    f[parallel[p]] = feq(c, parallel[p], rhoWall, u, v, w);
                for (int p = 0; p < nout; p++) {
                    const int q = outgoing[p];
                    const double cx = CS * XI[q * D];
                    const double cy = CS * XI[q * D + 1];
                    const double cz = D == 3 ? CS * XI[q * D + 2] : 0.0;
                    if (D == 2)
                        f[q] = f[c->OPP[q]] +
                               2 * rhoWall * c->W[q] * (cx * u + cy * v);
                    else
                        f[q] = f[c->OPP[q]] + 2 * rhoWall * c->W[q] *
                                                  (cx * u + cy * v + cz * w);
                }
            }
}

/* inward normal of the six block faces, as used by the extrapolation */
static int surface_face_code(int s) {
    switch (s) {
        case BS_LEFT: return 10;
        case BS_RIGHT: return 11;
        case BS_BOTTOM: return 20;
        case BS_TOP: return 21;
        case BS_BACK: return 30;
        case BS_FRONT: return 31;
    }
    return 0;
}

/* KerCutCellExtrapolPressure1ST[3D], boundary_kernel.inc:404-488, 894-1182.
 * Nodes of one face never read each other, so the order inside the face
 * range does not matter; the final rescale stops before the last
 * population (":484", ":1178"). */
static void bc_extrapol(orc_case* c, OBlock* b, const int r[6], int surface,
                        double rhoGiven) {
    const int Q = c->Q;
    const int scode = surface_face_code(surface);
    for (int k = r[4]; k < r[5]; k++)
        for (int j = r[2]; j < r[3]; j++)
            for (int i = r[0]; i < r[1]; i++) {
                size_t n = nidx(b, i, j, k);
                double* f = &b->f[n * Q];
                const int vg = b->vg[n];
                int ax[3], sg[3], inner;
                const int nf = vg_decode(vg, ax, sg, &inner);
                /* offset: a face copies from its inward neighbour whatever
                 * the surface; an edge or corner only when the surface is
                 * one of its faces */
                int off[3] = {0, 0, 0}, have = 0;
                if (nf == 1 && inner) {
                    off[ax[0]] = sg[0]; have = 1;
                } else if (nf >= 2 && inner) {
                    for (int a = 0; a < nf; a++)
                        if (scode == (ax[a] + 1) * 10 + (sg[a] > 0 ? 0 : 1)) {
                            off[ax[a]] = sg[a]; have = 1;
                        }
                }
                double rho = 0;
                for (int q = 0; q < Q; q++) {
                    int cv[3] = {XIi(c, q, 0), XIi(c, q, 1), XIi(c, q, 2)};
                    int unknown = 0;
                    if (nf >= 1 && inner) {
                        int allge = 1, nz = 0;
                        for (int a = 0; a < nf; a++) {
                            int cn = sg[a] * cv[ax[a]];
                            allge = allge && cn >= 0;
                            nz = nz || cn != 0;
                        }
                        unknown = nf == 1 ? (sg[0] * cv[ax[0]] > 0)
                                          : (allge && nz);
                    }
                    if (unknown && have) {
                        size_t src = nidx(b, i + off[0], j + off[1], k + off[2]);
                        f[q] = b->f[src * Q + q];
                    }
                    rho += f[q];
                }
                const double ratio = rhoGiven / rho;
                for (int q = 0; q < Q - 1; q++) f[q] *= ratio;
            }
}

/* KerCutCellPeriodic (boundary_kernel.inc:615-719) and
 * KerCutCellPeriodic3D (:1278-1666): every population of a boundary node
 * is copied from the node just outside the block along the normal of a
 * face of its geometry (its halo cell).  A face node copies whatever the
 * surface; an edge or corner node copies for each of its faces, in the
 * kernel's I, J, K order, whose surface is the boundary's (a later copy
 * overwrites an earlier one), except where the kernel tests another
 * surface (B4): at VG_IPKM_I the IP copy tests Right (:1392), at
 * VG_JMKP_I the KP copy tests Front (:1468), at VG_IPJMKM_I the IP copy
 * tests Top (:1550).  Halo cells are never written here, so the order
 * within the face range does not matter. */
static int fd_surface_test(int vg, int part_code) {
    /* the surface a part (face code 10 IP ... 31 KM) of geometry vg tests */
    if (vg == 10310 && part_code == 10) return BS_RIGHT;   /* VG_IPKM_I */
    if (vg == 21300 && part_code == 30) return BS_FRONT;   /* VG_JMKP_I */
    if (vg == 1021310 && part_code == 10) return BS_TOP;   /* VG_IPJMKM_I */
    switch (part_code) {
        case 10: return BS_LEFT;
        case 11: return BS_RIGHT;
        case 20: return BS_BOTTOM;
        case 21: return BS_TOP;
        case 30: return BS_BACK;
        case 31: return BS_FRONT;
    }
    return -100;
}

static void bc_fdperiodic(orc_case* c, OBlock* b, const int r[6], int surface) {
    const int Q = c->Q;
    for (int k = r[4]; k < r[5]; k++)
        for (int j = r[2]; j < r[3]; j++)
            for (int i = r[0]; i < r[1]; i++) {
                const size_t n = nidx(b, i, j, k);
                const int vg = b->vg[n];
                int ax[3], sg[3], inner;
                const int nf = vg_decode(vg, ax, sg, &inner);
                if (nf == 0 || !inner) continue;
                for (int a = 0; a < nf; a++) {
                    const int code = (ax[a] + 1) * 10 + (sg[a] > 0 ? 0 : 1);
                    if (nf > 1 && fd_surface_test(vg, code) != surface) continue;
                    int off[3] = {0, 0, 0};
                    off[ax[a]] = -sg[a]; /* the cell outside the face */
                    const size_t src = nidx(b, i + off[0], j + off[1], k + off[2]);
                    for (int q = 0; q < Q; q++) b->f[n * Q + q] = b->f[src * Q + q];
                }
            }
}

/* ImplementBoundary[3D] -> TreatBlockBoundary[3D], boundary.cpp:379-398,
 * boundary_wrapper.cpp:12-131 */
static void boundaries(orc_case* c) {
    for (int ib = 0; ib < c->nbc; ib++) {
        OBC* bc = &c->bc[ib];
        OBlock* b = &c->b[bc->block];
        int r[6];
        surface_range(c, b, bc->surface, r);
        if (bc->scheme == SCH_EQMDIFFUSEREFL)
            bc_eqm(c, b, r, bc->vars);
        else if (bc->scheme == SCH_EXTRAPOLPRESSURE1ST)
            bc_extrapol(c, b, r, bc->surface, bc->vars[0]);
        else if (bc->scheme == SCH_FDPERIODIC)
            bc_fdperiodic(c, b, r, bc->surface);
    }
}

/* ---- the swap scheme (Scheme_StreamCollision_Swap, 3D) -------------------
 * SwapStreamCollision (evolution.cpp:195-252) keeps one lattice, f, with
 * halos (scheme.cpp:102-105): macros -> collision in place -> body force
 * added in place -> halo transfer of f -> local swap -> swap stream ->
 * boundaries.  The velocity order puts opposites half a set apart
 * (opp(i) = i + 9 for i = 1..9 in D3Q19), which the swaps rely on (J.
 * Latt, "How to implement your DdQq dynamics with only q variables per
 * node", the reference's cited method). */

/* KerSwapCollideBGKIsothermal3D, model_kernel.inc:385-408: every node of
 * the whole range, whatever its type */
static void swap_collide(orc_case* c) {
    const int Q = c->Q;
    const double tau = c->tau, dt = c->dt;
    const double dtOvertauPlusdt = dt / (tau + 0.5 * dt);
    for (int ib = 0; ib < c->nb; ib++) {
        OBlock* b = &c->b[ib];
        OMP_FOR
        for (int k = 0; k < b->n[2]; k++)
            for (int j = 0; j < b->n[1]; j++)
                for (int i = 0; i < b->n[0]; i++) {
                    const size_t n = nidx(b, i, j, k);
                    const double rho = b->m[0][n], u = b->m[1][n], v = b->m[2][n],
                                 w = b->m[3][n];
                    double* f = &b->f[n * Q];
                    for (int q = 0; q < Q; q++) {
                        const double fe = feq(c, q, rho, u, v, w);
                        f[q] = fe + (1 - dtOvertauPlusdt) * (f[q] - fe);
                    }
                }
    }
}

/* KerSwapCalcBodyForce1ST3D, model_kernel.inc:543-568: f += F at Fluid and
 * MDPeriodic nodes (BodyForce_1st_Swap; BodyForce_None_Swap adds nothing) */
static void swap_body_force(orc_case* c) {
    if (c->force_type != 1) return;
    const int Q = c->Q, D = c->dim;
    const double CS = c->CS;
    const double* XI = c->XI;
    for (int ib = 0; ib < c->nb; ib++) {
        OBlock* b = &c->b[ib];
        OMP_FOR
        for (int k = 0; k < b->n[2]; k++)
            for (int j = 0; j < b->n[1]; j++)
                for (int i = 0; i < b->n[0]; i++) {
                    const size_t n = nidx(b, i, j, k);
                    const int vt = b->nt[n];
                    if (!(vt == VT_FLUID || vt == VT_MDPERIODIC)) continue;
                    for (int q = 0; q < Q; q++) {
                        double cf = 0;
                        for (int d = 0; d < D; d++) cf += CS * XI[q * D + d] * b->g[d];
                        const double F = c->W[q] * b->m[0][n] * cf;
                        b->f[n * Q + q] += F;
                    }
                }
    }
}

static int fluid_like(int vt) {
    return vt == VT_FLUID || vt == VT_MDPERIODIC || vt == VT_VIRTUALBOUNDARY;
}

static void swap_pair(double* a, double* b) {
    const double t = *a;
    *a = *b;
    *b = t;
}

/* KerLocalSwap3D, scheme_kernel.inc:959-1194, in the kernel's own
 * sequence: a fluid node exchanges f_i and f_{i+half}; a boundary node
 * (inner faces, edges and corners) sets f_i = f_opp(i) for the velocities
 * that point out of the block along every normal of its geometry (sign *
 * c <= 0 on each, not all zero), and swaps the pairs that are zero on
 * every normal (twice, once from each side: an identity) */
static void local_swap(orc_case* c) {
    const int Q = c->Q, half = (Q - 1) / 2;
    for (int ib = 0; ib < c->nb; ib++) {
        OBlock* b = &c->b[ib];
        OMP_FOR
        for (int k = 0; k < b->n[2]; k++)
            for (int j = 0; j < b->n[1]; j++)
                for (int i = 0; i < b->n[0]; i++) {
                    const size_t n = nidx(b, i, j, k);
                    const int vt = b->nt[n];
                    double* f = &b->f[n * Q];
                    if (fluid_like(vt)) {
                        for (int q = 1; q <= half; q++) swap_pair(&f[q], &f[q + half]);
                        continue;
                    }
                    if (vt == VT_IMMERSEDSOLID) continue;
                    int ax[3], sg[3], inner;
                    const int nf = vg_decode(b->vg[n], ax, sg, &inner);
                    if (nf == 0 || !inner) continue;
                    for (int q = 1; q < Q; q++) {
                        int allle = 1, allzero = 1;
                        for (int a = 0; a < nf; a++) {
                            const int cn = sg[a] * XIi(c, q, ax[a]);
                            allle = allle && cn <= 0;
                            allzero = allzero && cn == 0;
                        }
                        if (!allle) continue;
                        if (allzero) {
                            if (nf < 3) swap_pair(&f[q], &f[c->OPP[q]]);
                        } else {
                            f[q] = f[c->OPP[q]];
                        }
                    }
                }
    }
}

/* KerSwapStream3D, scheme_kernel.inc:1196-1398: a fluid node exchanges
 * f_{i+half}(x) with f_i(x + c_i) for i = 1..half; a boundary node only
 * where the opposite velocity points out of the block, or along it, on
 * every normal of its geometry (sign * c_opp <= 0 on each).  Each
 * exchange touches two values no other exchange touches, so the node
 * order does not matter. */
static void swap_stream(orc_case* c) {
    const int Q = c->Q, half = (Q - 1) / 2;
    for (int ib = 0; ib < c->nb; ib++) {
        OBlock* b = &c->b[ib];
        for (int k = 0; k < b->n[2]; k++)
            for (int j = 0; j < b->n[1]; j++)
                for (int i = 0; i < b->n[0]; i++) {
                    const size_t n = nidx(b, i, j, k);
                    const int vt = b->nt[n];
                    if (vt == VT_IMMERSEDSOLID) continue;
                    int ax[3] = {0, 0, 0}, sg[3] = {0, 0, 0}, inner = 0, nf = 0;
                    if (!fluid_like(vt)) {
                        nf = vg_decode(b->vg[n], ax, sg, &inner);
                        if (nf == 0 || !inner) continue;
                    }
                    for (int q = 1; q <= half; q++) {
                        int ok = 1;
                        for (int a = 0; a < nf; a++)
                            ok = ok && sg[a] * XIi(c, c->OPP[q], ax[a]) <= 0;
                        if (!ok) continue;
                        const size_t nb = nidx(b, i + XIi(c, q, 0), j + XIi(c, q, 1),
                                               k + XIi(c, q, 2));
                        swap_pair(&b->f[n * Q + q + half], &b->f[nb * Q + q]);
                    }
                }
    }
}

int orc_set_scheme(orc_case* c, int scheme) {
    if (scheme != SCHEME_STREAM_COLLISION && scheme != SCHEME_SWAP)
        return fail(c, "scheme %d is not restated", scheme);
    if (scheme == SCHEME_SWAP && (c->dim != 3 || c->Q != 19))
        return fail(c, "the swap scheme has D3Q19 kernels only (scheme_kernel.inc:957-1398)");
    c->scheme = scheme;
    return 0;
}

int orc_apply_boundaries(orc_case* c) {
    if (!c->prepared) return fail(c, "call orc_prepare first");
    boundaries(c);
    return 0;
}

int orc_step(orc_case* c, int nsteps) {
    if (!c->prepared) return fail(c, "call orc_prepare first");
    for (int s = 0; s < nsteps; s++) {
        if (c->scheme == SCHEME_SWAP) {
            orc_update_macros(c);
            swap_collide(c);
            swap_body_force(c);
            halo_transfer_field(c, 1);
            local_swap(c);
            swap_stream(c);
            boundaries(c);
            continue;
        }
        orc_update_macros(c);
        envelope_copy(c);
        body_force(c);
        collide(c);
        halo_transfer(c);
        stream(c);
        boundaries(c);
    }
    return 0;
}

/* ---- accessors ---------------------------------------------------------- */
int orc_get_f(orc_case* c, int block, double* out) {
    if (block < 0 || block >= c->nb) return fail(c, "undefined block");
    OBlock* b = &c->b[block];
    const int Q = c->Q;
    size_t p = 0;
    for (int q = 0; q < Q; q++)
        for (int k = 0; k < b->n[2]; k++)
            for (int j = 0; j < b->n[1]; j++)
                for (int i = 0; i < b->n[0]; i++)
                    out[p++] = b->f[nidx(b, i, j, k) * Q + q];
    return 0;
}

int orc_set_f(orc_case* c, int block, const double* in) {
    if (block < 0 || block >= c->nb) return fail(c, "undefined block");
    OBlock* b = &c->b[block];
    const int Q = c->Q;
    size_t p = 0;
    for (int q = 0; q < Q; q++)
        for (int k = 0; k < b->n[2]; k++)
            for (int j = 0; j < b->n[1]; j++)
                for (int i = 0; i < b->n[0]; i++)
                    b->f[nidx(b, i, j, k) * Q + q] = in[p++];
    return 0;
}

int orc_get_macro(orc_case* c, int block, int var, double* out) {
    if (block < 0 || block >= c->nb) return fail(c, "undefined block");
    if (var < 0 || var > c->dim) return fail(c, "bad macro variable");
    OBlock* b = &c->b[block];
    size_t p = 0;
    for (int k = 0; k < b->n[2]; k++)
        for (int j = 0; j < b->n[1]; j++)
            for (int i = 0; i < b->n[0]; i++)
                out[p++] = b->m[var][nidx(b, i, j, k)];
    return 0;
}

int orc_get_int(orc_case* c, int block, int which, int* out) {
    if (block < 0 || block >= c->nb) return fail(c, "undefined block");
    OBlock* b = &c->b[block];
    const int* src = which == 0 ? b->nt : b->vg;
    size_t p = 0;
    for (int k = 0; k < b->n[2]; k++)
        for (int j = 0; j < b->n[1]; j++)
            for (int i = 0; i < b->n[0]; i++) out[p++] = src[nidx(b, i, j, k)];
    return 0;
}

/* The total mass: no active reference function computes it (the uncompiled
 * CalcTotalMass, evolution_ToBeDeleted.cpp:282-294, is a plain loop).  A
 * plain running sum of 2e8 values near 1/3 drifts by about 1e-9 relative
 * (C3's 10.5M nodes: 0.035 at rest), far above the 1e-13 the checks ask
 * for, so the sum is compensated (Neumaier): within a few ulp of the exact
 * sum of the values. */
double orc_mass(orc_case* c) {
    double s = 0, comp = 0;
    const int Q = c->Q;
    for (int ib = 0; ib < c->nb; ib++) {
        OBlock* b = &c->b[ib];
        for (int k = 0; k < b->n[2]; k++)
            for (int j = 0; j < b->n[1]; j++)
                for (int i = 0; i < b->n[0]; i++) {
                    const double* f = &b->f[nidx(b, i, j, k) * Q];
                    for (int q = 0; q < Q; q++) {
                        const double v = f[q];
                        const double t = s + v;
                        if (fabs(s) >= fabs(v)) comp += (s - t) + v;
                        else comp += (v - t) + s;
                        s = t;
                    }
                }
    }
    return s + comp;
}

/* CalcResidualError (flowfield_wrapper.cpp:31-89): the squared difference
 * against the copy, the copy refreshed, then the sum of squares; the two
 * reductions start from zero (the reduction handle is reset by OPS when
 * its result is fetched, an assumption: OPS is not available). */
double orc_residual(orc_case* c, int check_period, double* per_var) {
    const int nv = 1 + c->dim;
    double diff[4] = {0, 0, 0, 0}, sum[4] = {0, 0, 0, 0};
    for (int v = 0; v < nv; v++)
        for (int ib = 0; ib < c->nb; ib++) {
            OBlock* b = &c->b[ib];
            for (int k = 0; k < b->n[2]; k++)
                for (int j = 0; j < b->n[1]; j++)
                    for (int i = 0; i < b->n[0]; i++) {
                        size_t n = nidx(b, i, j, k);
                        diff[v] = diff[v] + (b->m[v][n] - b->mc[v][n]) *
                                                (b->m[v][n] - b->mc[v][n]);
                    }
        }
    copy_current_macros(c);
    for (int v = 0; v < nv; v++)
        for (int ib = 0; ib < c->nb; ib++) {
            OBlock* b = &c->b[ib];
            for (int k = 0; k < b->n[2]; k++)
                for (int j = 0; j < b->n[1]; j++)
                    for (int i = 0; i < b->n[0]; i++) {
                        size_t n = nidx(b, i, j, k);
                        sum[v] = sum[v] + (b->m[v][n]) * (b->m[v][n]);
                    }
        }
    /* GetMaximumResidual, flowfield.cpp:148-158 */
    double maxRes = 0;
    for (int v = 0; v < nv; v++) {
        const double err = diff[v] / sum[v];
        double rel = err / (check_period * c->dt);
        if (per_var) per_var[v] = err;
        if (maxRes <= rel) maxRes = rel;
    }
    return maxRes;
}
