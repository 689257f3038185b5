// plan.hpp -- case description and the boundary "region programs" that the
// lattice kernels execute.
//
// A block is split into 3^d regions by position along each axis (interior,
// low face, high face).  Every reference construct that tags nodes or
// applies a boundary works on whole surface ranges (Block::RangeStart/End,
// block.cpp:14-97; SetBlockGeometryProperty flowfield_wrapper.cpp:291-623;
// SetBoundaryNodeType :625-669; TreatBlockBoundary boundary_wrapper.cpp),
// and every such range is a union of these regions.  So node type, geometry
// and the ordered boundary list are constant per region and the kernels need
// no per-node flag array: a node derives its region from its coordinates.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mplbhip {

constexpr int kMaxQ = 19;
constexpr int kMaxOps = 8;        // boundary applications per region
constexpr int kMaxBlocks = 4096;

// reference enumerations (values as declared there)
enum VertexType : int {
    VT_Fluid = 10, VT_Inlet = 11, VT_OutLet = 12, VT_MDPeriodic = 13,
    VT_Symmetry = 14, VT_FDPeriodic = 15, VT_Wall = 1000,
    VT_ImmersedSolid = -1, VT_ImmersedBoundary = -2, VT_VirtualBoundary = -3
};
enum BoundarySurface : int {
    BS_Left = 0, BS_Right = 1, BS_Top = 2, BS_Bottom = 3, BS_Front = 4,
    BS_Back = 5, BS_LeftBack = 50, BS_LeftFront = 40, BS_RightBack = 51,
    BS_RightFront = 41, BS_TopBack = 52, BS_TopFront = 42,
    BS_BottomBack = 53, BS_BottomFront = 43, BS_LeftTop = 20,
    BS_LeftBottom = 30, BS_RightTop = 21, BS_RightBottom = 31
};
enum BoundaryScheme : int {
    SCH_ExtrapolPressure1ST = 16, SCH_MDPeriodic = 18, SCH_FDPeriodic = 19,
    SCH_BounceBack = 20, SCH_EQMDiffuseRefl = 24, SCH_None = -1
};
enum VG : int { VG_Fluid = 0, VG_ImmersedSolid = -1 };

// ---- device-visible plain data ------------------------------------------
enum OpType : int32_t { OP_EQM = 1, OP_EXTRAPOL = 2, OP_FDPERIODIC = 3 };
// SchemeType (Src/scheme.h:69-72)
enum SchemeKind : int { SCHEME_StreamCollision = 10, SCHEME_Swap = 11 };
enum RegionKind : int32_t { RK_SKIP = 0, RK_FULL = 1, RK_MASKED = 2 };

struct BCOp {
    int32_t type;
    uint32_t m_in, m_out, m_par;  // EQMDiffuseRefl velocity classes
    uint32_t m_ext;               // extrapolation: populations copied
    int32_t off[3];               // extrapolation: inward neighbour offset;
                                  // FDPeriodic: the halo cell copied from
    int32_t nb_region;            // region of that neighbour
    int32_t nb_nops;              // neighbour ops applied before the read
    int32_t pad;
    double vars[3];               // wall velocity, or rho given in vars[0]
    // EQMDiffuseRefl in deviation form (g = f - w), the terms that depend
    // on the wall velocity only, for the reduced-storage bulk kernel:
    // rho_in = eq_win + sum_in g, rho_w = rho_in eq_k, then g = rho_w
    // eq_c - w (parallel) and g = g_opp + rho_w eq_c (outgoing)
    float eq_win;                 // sum of w over m_in
    float eq_k;                   // 2 / (1 - dRho - rhoPar)
    float eq_c[kMaxQ];            // parallel: w res(u); outgoing: 2 w c.u
};

struct RegionProg {
    int32_t kind;        // RegionKind
    int32_t force;       // Fluid or MDPeriodic: body force + shifted velocity
    uint32_t pull;       // populations streamed from the neighbours
    int32_t persist;     // element offset of this region's persist slots, -1 none
    int32_t ext[3];      // region extent (nodes) per axis, for persist indexing
    int32_t simple;      // pull (full or masked) + EQM ops only, no persisted
                         // state: the bulk kernel may run it inline
    int32_t vt, vg;      // node type and geometry of the region (the swap
                         // scheme's local swap and swap stream decode vg)
    int32_t nops;
    // faces of the region across which nothing streams, bits 2a (low
    // face of axis a) and 2a+1 (high face): a bulk-kernel lane moves the
    // source of a population that crosses one back inside the block (that
    // population is never streamed, and the row it then reads is one its
    // wave loads anyway).  Simple regions only.
    uint32_t adj;
    BCOp op[kMaxOps];
};

// ---- host-side case description -----------------------------------------
struct LinkDesc {
    int from, fsurf, to, tsurf, type;
    int rot;  // 0: reference link (same axis); 1: perpendicular faces, turned
};

// Geometry of a halo link: R (new = R old, a signed permutation) and the
// face axes/sides.  A reference link (rot = 0) is the identity into the
// halo opposite fsurf, whatever tsurf says (Field::CreateHalos uses fsurf
// only); a rotated link turns by the quarter rotation about the third axis
// that takes the from-face's outward normal to the to-face's inward one.
struct LinkMap {
    int R[3][3];
    int af, sf;  // from-face axis and outward side (+1 high, -1 low)
    int at, st;  // target halo axis and side of the target block
};
int face_axis(int surface, int* side);  // -1 for non-face surfaces
bool link_map(const LinkDesc& l, LinkMap& m);
// target-block node receiving from-block node p of the from-face plane
// (transverse halo rows included); nt = target block size
void link_target(const LinkMap& m, const int nt[3], const int p[3], int out[3]);
struct BCDesc {
    int block, surface, scheme, vtype, nvars;
    double vars[8];
};
struct BlockDef {
    int n[3];
};

struct Plan {
    int dim = 0, Q = 0;
    int lat = 0;  // lattice key: 2 D2Q9, 3 D3Q19, 15 D3Q15 (0: by dim)
    int scheme = SCHEME_StreamCollision;
    std::vector<BlockDef> blocks;
    std::vector<LinkDesc> links;
    std::vector<BCDesc> bcs;
    // outputs of build()
    int nregions = 0;                 // 9 (2D) or 27 (3D)
    std::vector<RegionProg> progs;    // blocks.size() * nregions
    std::vector<int> region_vt, region_vg;  // per block * region
    std::vector<long long> persist_nodes;   // per block*region (0 = none)
    long long persist_total = 0;      // nodes with persisted populations
    std::string error;

    bool build();
    // node type / geometry of node (i,j,k) of block b (host query)
    int node_type(int b, int i, int j, int k) const;
    int geometry(int b, int i, int j, int k) const;
    int region_of(int b, int i, int j, int k) const;
};

// lattice tables (reference order, Src/model.cpp:99-131)
extern const int kD2Q9[9][3];
extern const double kD2Q9W[9];
extern const int kD3Q19[19][3];
extern const double kD3Q19W[19];
extern const int kD2Q9Opp[9];
extern const int kD3Q19Opp[19];
// D3Q15, Src/model.cpp:133-150
extern const int kD3Q15[15][3];
extern const double kD3Q15W[15];
extern const int kD3Q15Opp[15];

// the tables of a lattice key (2 D2Q9, 3 D3Q19, 15 D3Q15)
struct LatticeTab {
    int Q;
    const int (*C)[3];
    const double* W;
    const int* OPP;
};
LatticeTab lattice_tab(int lat);

// helpers exposed for tests of the plan
void surface_range(int dim, const int n[3], int surface, int r[6]);
int vg_decode(int vg, int ax[3], int sg[3], int* inner);
bool stream_pulls(int vg, const int c[3]);
int dv_type(int vg, const int c[3]);  // 1 in, 2 out, 3 parallel, -1 invalid

}  // namespace mplbhip
