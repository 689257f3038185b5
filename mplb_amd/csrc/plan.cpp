// plan.cpp -- node typing, geometry and per-region boundary programs.
#include "plan.hpp"

#include <cmath>

#include <cstdio>
#include <cstring>

namespace mplbhip {

// Src/model.cpp:99-112 and :114-131 (opposites: FindReverseXi :82-97)
const int kD2Q9[9][3] = {{0, 0, 0},  {-1, 1, 0}, {-1, 0, 0},
                         {-1, -1, 0}, {0, -1, 0}, {1, -1, 0},
                         {1, 0, 0},  {1, 1, 0},  {0, 1, 0}};
const double kD2Q9W[9] = {4.0 / 9.0, 1.0 / 36.0, 1.0 / 9.0,
                          1.0 / 36.0, 1.0 / 9.0, 1.0 / 36.0,
                          1.0 / 9.0, 1.0 / 36.0, 1.0 / 9.0};
const int kD2Q9Opp[9] = {0, 5, 6, 7, 8, 1, 2, 3, 4};
const int kD3Q19[19][3] = {
    {0, 0, 0},  {-1, 0, 0},  {0, -1, 0}, {0, 0, -1}, {-1, -1, 0},
    {-1, 1, 0}, {-1, 0, -1}, {-1, 0, 1}, {0, -1, -1}, {0, -1, 1},
    {1, 0, 0},  {0, 1, 0},   {0, 0, 1},  {1, 1, 0},  {1, -1, 0},
    {1, 0, 1},  {1, 0, -1},  {0, 1, 1},  {0, 1, -1}};
const double kD3Q19W[19] = {
    1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0,
    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
    1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0,
    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
const int kD3Q19Opp[19] = {0,  10, 11, 12, 13, 14, 15, 16, 17, 18,
                           1,  2,  3,  4,  5,  6,  7,  8,  9};
// Src/model.cpp:133-150 (weights 2/9, 1/9, 1/72); opposites by the same
// first-match search as FindReverseXi
const int kD3Q15[15][3] = {
    {0, 0, 0},   {-1, 0, 0},  {0, -1, 0}, {0, 0, -1}, {-1, -1, -1},
    {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1}, {1, 0, 0},  {0, 1, 0},
    {0, 0, 1},   {1, 1, 1},   {1, 1, -1}, {1, -1, 1}, {1, -1, -1}};
const double kD3Q15W[15] = {2.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 72.0,
                            1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 9.0,  1.0 / 9.0,
                            1.0 / 9.0,  1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0};
const int kD3Q15Opp[15] = {0, 8, 9, 10, 11, 12, 13, 14, 1, 2, 3, 4, 5, 6, 7};

LatticeTab lattice_tab(int lat) {
    if (lat == 2) return LatticeTab{9, kD2Q9, kD2Q9W, kD2Q9Opp};
    if (lat == 15) return LatticeTab{15, kD3Q15, kD3Q15W, kD3Q15Opp};
    return LatticeTab{19, kD3Q19, kD3Q19W, kD3Q19Opp};
}

// Block::RangeStart/RangeEnd (block.cpp:14-97), including the "OP_3D"
// misspelling at block.cpp:35 that leaves TopBack/TopFront on the full y
// range.  r = {x0,x1,y0,y1,z0,z1}, half open.
void surface_range(int dim, const int n[3], int s, int r[6]) {
    for (int a = 0; a < 3; a++) {
        r[2 * a] = 0;
        r[2 * a + 1] = a < dim ? n[a] : 1;
    }
    const bool d3 = dim == 3;
    if (s == BS_Right || s == BS_RightTop || s == BS_RightBottom ||
        (d3 && (s == BS_RightFront || s == BS_RightBack)))
        r[0] = n[0] - 1;
    if (s == BS_Left || s == BS_LeftBottom || s == BS_LeftTop ||
        (d3 && (s == BS_LeftFront || s == BS_LeftBack)))
        r[1] = 1;
    if (s == BS_Top || s == BS_LeftTop || s == BS_RightTop) r[2] = n[1] - 1;
    if (s == BS_Bottom || s == BS_LeftBottom || s == BS_RightBottom ||
        (d3 && (s == BS_BottomBack || s == BS_BottomFront)))
        r[3] = 1;
    if (d3) {
        if (s == BS_Front || s == BS_LeftFront || s == BS_RightFront ||
            s == BS_TopFront || s == BS_BottomFront)
            r[4] = n[2] - 1;
        if (s == BS_Back || s == BS_LeftBack || s == BS_RightBack ||
            s == BS_TopBack || s == BS_BottomBack)
            r[5] = 1;
    }
}

static bool in_range(const int r[6], int i, int j, int k) {
    return i >= r[0] && i < r[1] && j >= r[2] && j < r[3] && k >= r[4] &&
           k < r[5];
}

// Geometry codes concatenate two-digit face codes (10 IP, 11 IM, 20 JP,
// 21 JM, 30 KP, 31 KM) plus, for edges and corners, a last digit 0 (_I) or
// 1 (_O) (flowfield_host_device.h:77-144).
int vg_decode(int vg, int ax[3], int sg[3], int* inner) {
    int nf, codes[3];
    if (vg >= 10 && vg < 100) {
        nf = 1;
        codes[0] = vg;
        *inner = 1;
    } else if (vg >= 10000 && vg < 100000) {
        nf = 2;
        *inner = (vg % 10) == 0;
        vg /= 10;
        codes[0] = vg / 100;
        codes[1] = vg % 100;
    } else if (vg >= 1000000 && vg < 10000000) {
        nf = 3;
        *inner = (vg % 10) == 0;
        vg /= 10;
        codes[0] = vg / 10000;
        codes[1] = (vg / 100) % 100;
        codes[2] = vg % 100;
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

// KerStream[3D] pull predicate at boundary nodes (scheme_kernel.inc:80-217,
// 1428-2049): sign*c <= 0 along every normal, all axes (_I) or any (_O).
bool stream_pulls(int vg, const int c[3]) {
    int ax[3], sg[3], inner;
    const int nf = vg_decode(vg, ax, sg, &inner);
    if (nf == 0) return false;
    bool all = true, any = false;
    for (int i = 0; i < nf; i++) {
        const bool ok = sg[i] * c[ax[i]] <= 0;
        all = all && ok;
        any = any || ok;
    }
    return (nf == 1 || inner) ? all : any;
}

// FindBdyDvType[3D] (boundary_host_device.h:47-183, 191-706) written with
// the components along the inward normals.  Corners keep the reference's
// naming: "outgoing" is the wall-pointing set there.
int dv_type(int vg, const int c[3]) {
    int ax[3], sg[3], inner;
    const int nf = vg_decode(vg, ax, sg, &inner);
    if (nf == 0) return -1;
    int cn[3] = {0, 0, 0};
    for (int i = 0; i < nf; i++) cn[i] = sg[i] * c[ax[i]];
    if (nf == 1) return cn[0] < 0 ? 1 : (cn[0] == 0 ? 3 : 2);
    if (nf == 2) {
        if ((cn[0] <= 0 && cn[1] < 0) || (cn[0] < 0 && cn[1] == 0)) return 1;
        if ((cn[0] >= 0 && cn[1] > 0) || (cn[0] > 0 && cn[1] == 0)) return 2;
        return 3;
    }
    if (c[0] == 0 && c[1] == 0 && c[2] == 0) return 3;
    if (cn[0] <= 0 && cn[1] <= 0 && cn[2] <= 0) return 2;
    if (cn[0] >= 0 && cn[1] >= 0 && cn[2] >= 0) return 1;
    return 3;
}

static int face_code_of_surface(int s) {
    switch (s) {
        case BS_Left: return 10;
        case BS_Right: return 11;
        case BS_Bottom: return 20;
        case BS_Top: return 21;
        case BS_Back: return 30;
        case BS_Front: return 31;
        default: return 0;
    }
}

// The surface the FDPeriodic kernels test for the face part `code` (10 IP
// ... 31 KM) of geometry vg: the face's own surface, except where
// KerCutCellPeriodic3D tests another one (boundary_kernel.inc:1392 IP of
// VG_IPKM_I tests Right, :1468 KP of VG_JMKP_I tests Front, :1550 IP of
// VG_IPJMKM_I tests Top).
static int fd_surface_test(int vg, int code) {
    if (vg == 10310 && code == 10) return BS_Right;
    if (vg == 21300 && code == 30) return BS_Front;
    if (vg == 1021310 && code == 10) return BS_Top;
    switch (code) {
        case 10: return BS_Left;
        case 11: return BS_Right;
        case 20: return BS_Bottom;
        case 21: return BS_Top;
        case 30: return BS_Back;
        case 31: return BS_Front;
        default: return -100;
    }
}

// SetBlockGeometryProperty (flowfield_wrapper.cpp:291-623) at one node of
// the whole range (halo slabs are outside it and not queried).
int Plan::geometry(int b, int i, int j, int k) const {
    const int* n = blocks[b].n;
    const int nx = n[0], ny = n[1], nz = dim == 3 ? n[2] : 1;
    int vg = 0;
    auto set = [&](int x0, int x1, int y0, int y1, int z0, int z1, int v) {
        const int r[6] = {x0, x1, y0, y1, z0, z1};
        if (in_range(r, i, j, k)) vg = v;
    };
    set(1, nx - 1, 1, ny - 1, dim == 3 ? 1 : 0, dim == 3 ? nz - 1 : 1,
        VG_Fluid);
    int r[6];
    const int faces[6][2] = {{BS_Bottom, 20}, {BS_Top, 21}, {BS_Left, 10},
                             {BS_Right, 11},  {BS_Back, 30}, {BS_Front, 31}};
    for (int f = 0; f < (dim == 3 ? 6 : 4); f++) {
        surface_range(dim, n, faces[f][0], r);
        if (in_range(r, i, j, k)) vg = faces[f][1];
    }
    if (dim == 2) {
        set(0, 1, 0, 1, 0, 1, 10200);
        set(0, 1, ny - 1, ny, 0, 1, 10210);
        set(nx - 1, nx, ny - 1, ny, 0, 1, 11210);
        set(nx - 1, nx, 0, 1, 0, 1, 11200);
        return vg;
    }
    set(0, 1, 0, 1, 0, nz, 10200);
    set(0, 1, ny - 1, ny, 0, nz, 10210);
    set(nx - 1, nx, ny - 1, ny, 0, nz, 11210);
    set(nx - 1, nx, 0, 1, 0, nz, 11200);
    set(0, 1, 0, ny, 0, 1, 10300);
    set(0, 1, 0, ny, nz - 1, nz, 10310);
    set(nx - 1, nx, 0, ny, nz - 1, nz, 11310);
    set(nx - 1, nx, 0, ny, 0, 1, 11300);
    set(0, nx, 0, 1, 0, 1, 20300);
    set(0, nx, 0, 1, nz - 1, nz, 20310);
    set(0, nx, ny - 1, ny, nz - 1, nz, 21310);
    set(0, nx, ny - 1, ny, 0, 1, 21300);
    set(0, 1, 0, 1, 0, 1, 1020300);
    set(0, 1, 0, 1, nz - 1, nz, 1020310);
    set(0, 1, ny - 1, ny, 0, 1, 1021300);
    set(0, 1, ny - 1, ny, nz - 1, nz, 1021310);
    set(nx - 1, nx, 0, 1, 0, 1, 1120300);
    set(nx - 1, nx, 0, 1, nz - 1, nz, 1120310);
    set(nx - 1, nx, ny - 1, ny, 0, 1, 1121300);
    set(nx - 1, nx, ny - 1, ny, nz - 1, nz, 1121310);
    return vg;
}

// SetBulkandHaloNodesType + SetBoundaryNodeType (flowfield_wrapper.cpp:
// 625-788) at one node of the whole range.  Envelope nodes no boundary
// covers keep the zero the storage was created with.
int Plan::node_type(int b, int i, int j, int k) const {
    const int* n = blocks[b].n;
    int vt = 0;
    const bool bulk = i >= 1 && i < n[0] - 1 && j >= 1 && j < n[1] - 1 &&
                      (dim == 2 || (k >= 1 && k < n[2] - 1));
    if (bulk) vt = VT_Fluid;
    for (const BCDesc& bc : bcs) {
        if (bc.block != b) continue;
        int r[6];
        surface_range(dim, n, bc.surface, r);
        if (bc.vtype == VT_VirtualBoundary) {
            if (bc.surface == BS_Left || bc.surface == BS_Right) {
                r[2]++; r[3]--;
                if (dim == 3) { r[4]++; r[5]--; }
            }
            if (bc.surface == BS_Top || bc.surface == BS_Bottom) {
                r[0]++; r[1]--;
                if (dim == 3) { r[4]++; r[5]--; }
            }
            if (dim == 3 && (bc.surface == BS_Front || bc.surface == BS_Back)) {
                r[0]++; r[1]--; r[2]++; r[3]--;
            }
        }
        if (in_range(r, i, j, k)) vt = bc.vtype;
    }
    return vt;
}

int Plan::region_of(int b, int i, int j, int k) const {
    const int* n = blocks[b].n;
    auto pos = [](int x, int nn) { return x == 0 ? 1 : (x == nn - 1 ? 2 : 0); };
    int r = pos(i, n[0]) + 3 * pos(j, n[1]);
    if (dim == 3) r += 9 * pos(k, n[2]);
    return r;
}

static void region_rep(int dim, const int n[3], int r, int p[3]) {
    int code[3] = {r % 3, (r / 3) % 3, r / 9};
    for (int a = 0; a < 3; a++) {
        if (a >= dim) { p[a] = 0; continue; }
        p[a] = code[a] == 0 ? 1 : (code[a] == 1 ? 0 : n[a] - 1);
    }
}

bool Plan::build() {
    const LatticeTab L = lattice_tab(lat ? lat : dim);
    const int (*C)[3] = L.C;
    const int* OPP = L.OPP;
    Q = L.Q;
    nregions = dim == 2 ? 9 : 27;
    const int nb = (int)blocks.size();
    const uint32_t all = (Q == 32) ? 0xffffffffu : ((1u << Q) - 1u);
    progs.assign((size_t)nb * nregions, RegionProg{});
    region_vt.assign((size_t)nb * nregions, 0);
    region_vg.assign((size_t)nb * nregions, 0);
    persist_nodes.assign((size_t)nb * nregions, 0);
    for (int b = 0; b < nb; b++)
        for (int a = 0; a < dim; a++)
            if (blocks[b].n[a] < 3) {
                error = "every block axis needs at least 3 nodes";
                return false;
            }
    for (const BCDesc& bc : bcs) {
        if (bc.scheme == SCH_BounceBack) {
            error = "boundary scheme BounceBack not supported (the reference's kernel is "
                    "commented out, boundary_kernel.inc:490-532)";
            return false;
        }
        if (bc.scheme == SCH_FDPeriodic && scheme != SCHEME_Swap) {
            // KerCutCellPeriodic[3D] copies f from the halo cells, which only
            // the swap scheme exchanges (scheme.cpp:96-105): under
            // stream-collision they hold the zeros of a new dat
            error = "FDPeriodic copies f's halo, which the stream-collision scheme never "
                    "exchanges (only fStage's, scheme.cpp:96-99): use it with "
                    "Scheme_StreamCollision_Swap";
            return false;
        }
        if (bc.scheme == SCH_ExtrapolPressure1ST && scheme == SCHEME_Swap) {
            error = "ExtrapolPressure1ST under the swap scheme is not supported (it reads a "
                    "neighbour that other boundary nodes rewrite in place)";
            return false;
        }
    }
    // per-region op lists with their global definition order
    std::vector<std::vector<int>> op_order((size_t)nb * nregions);
    for (int b = 0; b < nb; b++) {
        const int* n = blocks[b].n;
        for (int r = 0; r < nregions; r++) {
            int p[3];
            region_rep(dim, n, r, p);
            RegionProg& P = progs[(size_t)b * nregions + r];
            const int vt = node_type(b, p[0], p[1], p[2]);
            const int vg = geometry(b, p[0], p[1], p[2]);
            region_vt[(size_t)b * nregions + r] = vt;
            region_vg[(size_t)b * nregions + r] = vg;
            P.vt = vt;
            P.vg = vg;
            P.force = (vt == VT_Fluid || vt == VT_MDPeriodic) ? 1 : 0;
            P.persist = -1;
            int code[3] = {r % 3, (r / 3) % 3, r / 9};
            for (int a = 0; a < 3; a++)
                P.ext[a] = (a < dim && code[a] == 0) ? n[a] - 2 : 1;
            if (vt == VT_ImmersedSolid) {
                P.kind = RK_SKIP;
                P.pull = 0;
            } else if (vt == VT_Fluid || vt == VT_MDPeriodic ||
                       vt == VT_VirtualBoundary) {
                P.kind = RK_FULL;
                P.pull = all;
            } else {
                P.kind = RK_MASKED;
                uint32_t m = 1u;  // the rest population always streams
                for (int q = 1; q < Q; q++)
                    if (stream_pulls(vg, C[q])) m |= 1u << q;
                P.pull = m;
            }
            // boundary applications covering this region, in order
            for (int ib = 0; ib < (int)bcs.size(); ib++) {
                const BCDesc& bc = bcs[ib];
                if (bc.block != b) continue;
                if (bc.scheme != SCH_EQMDiffuseRefl &&
                    bc.scheme != SCH_ExtrapolPressure1ST && bc.scheme != SCH_FDPeriodic)
                    continue;
                int rr[6];
                surface_range(dim, n, bc.surface, rr);
                if (!in_range(rr, p[0], p[1], p[2])) continue;
                if (P.kind == RK_SKIP) continue;  // never happens on envelopes
                if (P.nops >= kMaxOps) {
                    error = "too many boundary applications on one region";
                    return false;
                }
                if (bc.scheme == SCH_FDPeriodic) {
                    // KerCutCellPeriodic[3D] (boundary_kernel.inc:615-719,
                    // 1278-1666): one copy of every population per face
                    // part of the geometry whose surface the kernel tests,
                    // in its I, J, K order (the last copy wins)
                    int ax[3], sg[3], inner;
                    const int nf = vg_decode(vg, ax, sg, &inner);
                    if (nf == 0 || !inner) continue;
                    for (int a = 0; a < nf; a++) {
                        const int code = (ax[a] + 1) * 10 + (sg[a] > 0 ? 0 : 1);
                        if (nf > 1 && fd_surface_test(vg, code) != bc.surface) continue;
                        if (P.nops >= kMaxOps) {
                            error = "too many boundary applications on one region";
                            return false;
                        }
                        BCOp& op = P.op[P.nops++];
                        std::memset(&op, 0, sizeof op);
                        op_order[(size_t)b * nregions + r].push_back(ib);
                        op.type = OP_FDPERIODIC;
                        op.off[ax[a]] = -sg[a];  // the cell outside that face
                    }
                    continue;
                }
                BCOp& op = P.op[P.nops++];
                std::memset(&op, 0, sizeof op);
                op_order[(size_t)b * nregions + r].push_back(ib);
                if (bc.scheme == SCH_EQMDiffuseRefl) {
                    op.type = OP_EQM;
                    for (int q = 0; q < Q; q++) {
                        const int t = dv_type(vg, C[q]);
                        if (t == 1) op.m_in |= 1u << q;
                        if (t == 2) op.m_out |= 1u << q;
                        if (t == 3) op.m_par |= 1u << q;
                    }
                    for (int d = 0; d < 3; d++)
                        op.vars[d] = d < bc.nvars ? bc.vars[d] : 0.0;
                } else {
                    op.type = OP_EXTRAPOL;
                    op.vars[0] = bc.nvars > 0 ? bc.vars[0] : 1.0;
                    int ax[3], sg[3], inner;
                    const int nf = vg_decode(vg, ax, sg, &inner);
                    const int scode = face_code_of_surface(bc.surface);
                    int off[3] = {0, 0, 0};
                    bool have = false;
                    if (nf == 1 && inner) {
                        off[ax[0]] = sg[0];
                        have = true;
                    } else if (nf >= 2 && inner) {
                        for (int a = 0; a < nf; a++)
                            if (scode == (ax[a] + 1) * 10 + (sg[a] > 0 ? 0 : 1)) {
                                off[ax[a]] = sg[a];
                                have = true;
                            }
                    }
                    if (have && nf >= 1 && inner) {
                        for (int q = 0; q < Q; q++) {
                            bool allge = true, nz = false;
                            for (int a = 0; a < nf; a++) {
                                const int cn = sg[a] * C[q][ax[a]];
                                allge = allge && cn >= 0;
                                nz = nz || cn != 0;
                            }
                            const bool unknown =
                                nf == 1 ? (sg[0] * C[q][ax[0]] > 0) : (allge && nz);
                            if (unknown) op.m_ext |= 1u << q;
                        }
                    }
                    for (int a = 0; a < 3; a++) op.off[a] = off[a];
                    op.nb_region = region_of(b, p[0] + off[0], p[1] + off[1],
                                             p[2] + off[2]);
                }
            }
        }
    }
    // resolve extrapolation neighbours: how many of the neighbour region's
    // own applications precede the read (definition order across the case)
    for (int b = 0; b < nb; b++)
        for (int r = 0; r < nregions; r++) {
            RegionProg& P = progs[(size_t)b * nregions + r];
            for (int o = 0; o < P.nops; o++) {
                BCOp& op = P.op[o];
                if (op.type != OP_EXTRAPOL) continue;
                const int my_order = op_order[(size_t)b * nregions + r][o];
                const auto& nbo = op_order[(size_t)b * nregions + op.nb_region];
                const RegionProg& N = progs[(size_t)b * nregions + op.nb_region];
                int cnt = 0;
                for (int k = 0; k < N.nops; k++)
                    if (nbo[k] < my_order) cnt = k + 1;
                for (int k = 0; k < cnt; k++)
                    if (N.op[k].type == OP_EXTRAPOL) {
                        error = "an extrapolated boundary reads a neighbour that "
                                "is itself extrapolated earlier: unsupported";
                        return false;
                    }
                if (N.kind == RK_SKIP) {
                    error = "an extrapolated boundary reads a solid node";
                    return false;
                }
                op.nb_nops = cnt;
            }
        }
    // Which populations must survive from the previous step?  A population
    // that is not streamed keeps its old value (scheme_kernel.inc:1417-1428)
    // and is needed when it is read before it is written, by the region's own
    // boundary sequence, by a neighbour's extrapolation, or simply because no
    // one overwrites it (it is part of the state).
    std::vector<uint32_t> need((size_t)nb * nregions, 0u);
    auto simulate = [&](const RegionProg& P, int upto, uint32_t* written) {
        uint32_t w = P.pull, needs = 0;
        for (int o = 0; o < upto; o++) {
            const BCOp& op = P.op[o];
            if (op.type == OP_EQM) {
                needs |= op.m_in & ~w;
                w |= op.m_par;
                for (int q = 0; q < Q; q++)
                    if (op.m_out >> q & 1u) {
                        if (!(w >> OPP[q] & 1u)) needs |= 1u << OPP[q];
                        w |= 1u << q;
                    }
            } else if (op.type == OP_FDPERIODIC) {
                w = all;  // every population copied from outside the block
            } else {
                w |= op.m_ext;
                needs |= all & ~w;  // the density sums every population
                w = all;
            }
        }
        *written = w;
        return needs;
    };
    for (int b = 0; b < nb; b++)
        for (int r = 0; r < nregions; r++) {
            const RegionProg& P = progs[(size_t)b * nregions + r];
            if (P.kind == RK_SKIP) continue;
            uint32_t w;
            uint32_t needs = simulate(P, P.nops, &w);
            needs |= all & ~w;
            need[(size_t)b * nregions + r] |= needs;
            for (int o = 0; o < P.nops; o++) {
                const BCOp& op = P.op[o];
                if (op.type != OP_EXTRAPOL) continue;
                const RegionProg& N = progs[(size_t)b * nregions + op.nb_region];
                uint32_t wn;
                simulate(N, op.nb_nops, &wn);
                need[(size_t)b * nregions + op.nb_region] |= op.m_ext & ~wn;
            }
        }
    persist_total = 0;
    for (int b = 0; b < nb; b++)
        for (int r = 0; r < nregions; r++) {
            RegionProg& P = progs[(size_t)b * nregions + r];
            if (need[(size_t)b * nregions + r] == 0u) continue;
            const long long cnt = (long long)P.ext[0] * P.ext[1] * P.ext[2];
            P.persist = (int32_t)persist_total;
            persist_nodes[(size_t)b * nregions + r] = cnt;
            persist_total += cnt;
        }
    if (persist_total > (1LL << 30)) {
        error = "persisted boundary state too large";
        return false;
    }
    const double cs = std::sqrt(3.0);
    for (int b = 0; b < nb; b++)
        for (int r = 0; r < nregions; r++) {
            RegionProg& P = progs[(size_t)b * nregions + r];
            bool eqm_only = true;
            for (int o = 0; o < P.nops; o++) eqm_only = eqm_only && P.op[o].type == OP_EQM;
            P.simple = (P.kind == RK_FULL || P.kind == RK_MASKED) && P.persist < 0 && eqm_only;
            for (int o = 0; o < P.nops; o++) {
                BCOp& op = P.op[o];
                if (op.type != OP_EQM) continue;
                double win = 0.0, drho = 0.0, rpar = 0.0;
                const double u[3] = {op.vars[0], op.vars[1], op.vars[2]};
                const double u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
                for (int q = 0; q < Q; q++) {
                    const double cu = cs * (C[q][0] * u[0] + C[q][1] * u[1] + C[q][2] * u[2]);
                    op.eq_c[q] = 0.f;
                    if (op.m_in >> q & 1u) {
                        win += L.W[q];
                    } else if (op.m_out >> q & 1u) {
                        drho += 2.0 * L.W[q] * cu;
                        op.eq_c[q] = (float)(2.0 * L.W[q] * cu);
                    } else if (op.m_par >> q & 1u) {
                        const double f1 = L.W[q] * (1.0 + cu + 0.5 * (cu * cu - u2));
                        rpar += f1;
                        op.eq_c[q] = (float)f1;
                    }
                }
                op.eq_win = (float)win;
                op.eq_k = (float)(2.0 / (1.0 - drho - rpar));
            }
            // faces the region does not stream across: the population that
            // points straight in from such a face is not pulled
            P.adj = 0u;
            if (!P.simple || P.kind != RK_MASKED) continue;
            int code[3] = {r % 3, (r / 3) % 3, r / 9};
            for (int a = 0; a < dim; a++) {
                if (code[a] == 0) continue;
                const int in = code[a] == 1 ? 1 : -1;  // c_a of a source outside
                for (int q = 1; q < Q; q++) {
                    bool unit = C[q][a] == in;
                    for (int d = 0; d < 3; d++) unit = unit && (d == a || C[q][d] == 0);
                    if (unit && !(P.pull >> q & 1u)) P.adj |= 1u << (2 * a + (code[a] - 1));
                }
            }
            // a lane shifts every population crossing a flagged face: each
            // such population must be one the region does not stream
            for (int q = 0; q < Q; q++) {
                bool crosses = false;
                for (int a = 0; a < dim; a++) {
                    if ((P.adj >> (2 * a) & 1u) && C[q][a] > 0) crosses = true;
                    if ((P.adj >> (2 * a + 1) & 1u) && C[q][a] < 0) crosses = true;
                }
                if (crosses && (P.pull >> q & 1u)) {
                    P.simple = 0;  // the boundary kernel runs it
                    P.adj = 0u;
                    break;
                }
            }
        }
    return true;
}

// ---- halo links ------------------------------------------------------------
int face_axis(int s, int* side) {
    *side = 0;
    switch (s) {
        case BS_Left: *side = -1; return 0;
        case BS_Right: *side = 1; return 0;
        case BS_Bottom: *side = -1; return 1;
        case BS_Top: *side = 1; return 1;
        case BS_Back: *side = -1; return 2;
        case BS_Front: *side = 1; return 2;
        default: return -1;
    }
}

bool link_map(const LinkDesc& l, LinkMap& m) {
    for (int j = 0; j < 3; j++)
        for (int k = 0; k < 3; k++) m.R[j][k] = 0;
    m.af = face_axis(l.fsurf, &m.sf);
    if (m.af < 0) return false;
    if (!l.rot) {
        for (int k = 0; k < 3; k++) m.R[k][k] = 1;
        m.at = m.af;
        m.st = -m.sf;
        return true;
    }
    m.at = face_axis(l.tsurf, &m.st);
    if (m.at < 0 || m.at == m.af) return false;
    const int a3 = 3 - m.af - m.at;
    const int sigma = -m.sf * m.st;  // R (sf e_af) = -st e_at
    m.R[m.at][m.af] = sigma;
    m.R[m.af][m.at] = -sigma;        // det R = +1
    m.R[a3][a3] = 1;
    return true;
}

void link_target(const LinkMap& m, const int nt[3], const int p[3], int out[3]) {
    for (int k = 0; k < 3; k++) {
        if (k == m.af) {
            out[m.at] = m.st > 0 ? nt[m.at] : -1;
            continue;
        }
        for (int j = 0; j < 3; j++) {
            if (m.R[j][k] == 1) out[j] = p[k];
            if (m.R[j][k] == -1) out[j] = nt[j] - 1 - p[k];
        }
    }
}

}  // namespace mplbhip
