// hilemms.cpp -- the HiLeMMS case API (DefineCase ... Iterate) on top of
// libmplbhip's C ABI.  One case, one component, one GPU (MPLB_DEVICE,
// default 0); population storage from MPLB_STORAGE = f64 (default, the
// bit-exact path), f32 or f16.
//
// Errors keep the reference's behaviour: a message through ops_printf and
// an abort (Src/*.cpp use ops_printf + assert).
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/mplbhip.h"
#include "hdf5_io.h"
#include "mplb.h"

// ---------------------------------------------------------------------------
// lattice globals of model.h
int NUMXI{9};
int LATTDIM{2};
Real CS{1};
Real* XI{nullptr};
Real XIMAXVALUE{1};
Real* WEIGHTS{nullptr};
int NUMCOMPONENTS{1};
int* OPP{nullptr};

namespace {

[[noreturn]] void die(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vfprintf(stderr, fmt, ap);
    va_end(ap);
    std::fprintf(stderr, "\n");
    std::fflush(stderr);
    std::abort();
}

struct State {
    std::string caseName;
    int spaceDim{0};
    bool transient{false};
    BlockGroup blocks;
    std::map<int, std::vector<std::vector<Real>>> coordinates;
    Real meshSize{0};
    std::map<int, Component> components;
    RealField f{"f"}, fStage{"fStage"}, coordinateXYZ{"CoordinateXYZ"};
    RealFieldGroup macroVars, macroVarsCopy, macroBodyforce;
    IntFieldGroup nodeType;
    IntField geometryProperty{"GeometryProperty"};
    std::map<int, Real> residualError;
    std::vector<BlockBoundary> boundaries;
    std::vector<std::array<int, 5>> connections;
    SchemeType scheme{Scheme_StreamCollision};
    Real dt{1};
    mplbhip_ctx* ctx{nullptr};
    bool partitioned{false};
    bool forceMacros{false};
    SizeType restartStep{0};  // DefineComponents(..., timeStep): restart from files
};

State& S() {
    static State s;
    return s;
}

void ck(int rc, const char* what) {
    if (rc != 0)
        die("Error! %s failed: %s", what,
            S().ctx ? mplbhip_last_error(S().ctx) : "no context");
}

int storage_from_env() {
    const char* s = std::getenv("MPLB_STORAGE");
    if (!s || !std::strcmp(s, "f64")) return MPLBHIP_STORE_F64;
    if (!std::strcmp(s, "f32")) return MPLBHIP_STORE_F32;
    if (!std::strcmp(s, "f16")) return MPLBHIP_STORE_F16;
    die("Error! MPLB_STORAGE must be f64, f32 or f16 (got %s)", s);
}

const Component& the_component() {
    if (S().components.size() != 1)
        die("Error! exactly one component is supported by this build");
    return S().components.begin()->second;
}

// copy the interior of a scalar host mirror into a dense [k][j][i] array
std::vector<double> interior(ops_dat d) {
    std::vector<double> out((size_t)d->size[0] * d->size[1] * d->size[2]);
    const double* p = reinterpret_cast<const double*>(d->data.data());
    size_t n = 0;
    for (int k = 0; k < d->size[2]; k++)
        for (int j = 0; j < d->size[1]; j++)
            for (int i = 0; i < d->size[0]; i++) out[n++] = p[d->offset(i, j, k)];
    return out;
}
void set_interior(ops_dat d, const std::vector<double>& v) {
    double* p = reinterpret_cast<double*>(d->data.data());
    size_t n = 0;
    for (int k = 0; k < d->size[2]; k++)
        for (int j = 0; j < d->size[1]; j++)
            for (int i = 0; i < d->size[0]; i++) p[d->offset(i, j, k)] = v[n++];
    d->dirty = false;
}

int var_slot(const Component& c, int varId) {
    for (const auto& mv : c.macroVars) {
        if (mv.second.id != varId) continue;
        switch (mv.first) {
            case Variable_Rho: return 0;
            case Variable_U: case Variable_U_Force: return 1;
            case Variable_V: case Variable_V_Force: return 2;
            case Variable_W: case Variable_W_Force: return 3;
            default: return -1;
        }
    }
    return -1;
}

// checkpoint files (Src/flowfield.cpp:118-137, Src/field.h:180-205) are
// written unless MPLB_HDF5=0 (benchmarks that do not want them)
bool hdf5_enabled() {
    const char* e = std::getenv("MPLB_HDF5");
    return !e || std::atoi(e) != 0;
}

template <typename T>
void write_field(Field<T>& fld, SizeType step) {
    for (const auto& idBlock : S().blocks) {
        if (!fld.Has(idBlock.first)) continue;
        const std::string& name = idBlock.second.Name();
        hilemms_h5::write_dat(hilemms_h5::file_name(S().caseName, name, (long long)step), name,
                              idBlock.first, S().spaceDim, fld.at(idBlock.first));
    }
}

template <typename T>
void read_field(Field<T>& fld, SizeType step) {
    for (const auto& idBlock : S().blocks) {
        const std::string& name = idBlock.second.Name();
        hilemms_h5::read_dat(hilemms_h5::file_name(S().caseName, name, (long long)step), name,
                             S().spaceDim, fld.at(idBlock.first));
    }
}

// the host mirror of f (AoS with its halo, as the reference's field f),
// created when a checkpoint or a restart first needs it
void ensure_f_mirror() {
    State& s = S();
    if (!s.blocks.empty() && s.f.Has(s.blocks.begin()->first)) return;
    s.f.SetDataDim(NUMXI);
    s.f.CreateFieldFromScratch(s.blocks);
}

void download_macros() {
    const Component& c = the_component();
    for (auto& idField : S().macroVars) {
        const int slot = var_slot(c, idField.first);
        if (slot < 0 || slot > S().spaceDim) continue;
        for (const auto& idBlock : S().blocks) {
            ops_dat d = idField.second.at(idBlock.first);
            std::vector<double> v((size_t)d->size[0] * d->size[1] * d->size[2]);
            ck(mplbhip_get_macro(S().ctx, idBlock.first, slot, v.data()), "get_macro");
            set_interior(d, v);
        }
    }
}

// the acceleration of one block from its g_MacroBodyforce field, which must
// be uniform: libmplbhip keeps one per block
std::array<double, 3> block_accel(ops_dat d, int block) {
    const double* p = reinterpret_cast<const double*>(d->data.data());
    std::array<double, 3> g{0, 0, 0};
    for (int a = 0; a < S().spaceDim; a++) g[a] = p[d->offset(0, 0, 0) + a];
    for (int k = 0; k < d->size[2]; k++)
        for (int j = 0; j < d->size[1]; j++)
            for (int i = 0; i < d->size[0]; i++)
                for (int a = 0; a < S().spaceDim; a++)
                    if (p[d->offset(i, j, k) + a] != g[a])
                        die("Error! the body force must be uniform within a "
                            "block (block %d)", block);
    return g;
}

// accelerations in force on the device, per block
std::map<int, std::array<double, 3>>& uploaded_accel() {
    static std::map<int, std::array<double, 3>> m;
    return m;
}

// true when the app's hook has written an acceleration that differs from
// the one in force (a rewrite of the same values only clears the flag)
bool force_changed() {
    if (S().macroBodyforce.empty()) return false;
    RealField& F = S().macroBodyforce.begin()->second;
    bool changed = false;
    for (const auto& idBlock : S().blocks) {
        ops_dat d = F.at(idBlock.first);
        if (!d->dirty) continue;
        const auto g = block_accel(d, idBlock.first);
        auto it = uploaded_accel().find(idBlock.first);
        if (it != uploaded_accel().end() && it->second == g)
            d->dirty = false;
        else
            changed = true;
    }
    return changed;
}

void upload_force_if_dirty() {
    if (S().macroBodyforce.empty()) return;
    RealField& F = S().macroBodyforce.begin()->second;
    for (const auto& idBlock : S().blocks) {
        ops_dat d = F.at(idBlock.first);
        if (!d->dirty) continue;
        const auto g = block_accel(d, idBlock.first);
        auto it = uploaded_accel().find(idBlock.first);
        if (it == uploaded_accel().end() || it->second != g) {
            ck(mplbhip_set_accel(S().ctx, idBlock.first, g.data()), "set_accel");
            uploaded_accel()[idBlock.first] = g;
        }
        d->dirty = false;
    }
}

// n device steps from time step iter.  The reference calls the App's
// UpdateMacroscopicBodyForce(iter * dt) before every step
// (evolution.cpp:137-151); so does this loop, on the host, at every step's
// time.  The device applies one acceleration per block to a step (its
// velocity shift and its source term alike), taken from the hook's first
// call of this Iterate; a hook that changes it at a later step stops the
// run with an error instead of being ignored.  (The Apps' hooks are empty.)
// criteria < 0: the transient form; the macros of check steps are refreshed
// on the device.  Returns the steps taken (fewer than n only when a check
// step met criteria); *res the last residual or -1.
long long advance(SizeType iter, SizeType n, double criteria, SizeType period, double* res,
                  bool first) {
    for (SizeType k = 0; k < n; k++) {
        UpdateMacroscopicBodyForce((iter + k) * TimeStep());
        if (first && k == 0) {
            upload_force_if_dirty();
        } else if (force_changed()) {
            die("Error! UpdateMacroscopicBodyForce changed the body force at time step %llu "
                "inside Iterate: the device loop applies the acceleration set before it; a "
                "time-dependent force is not supported (call Iterate per force value)",
                (unsigned long long)(iter + k));
        }
    }
    long long done = 0;
    *res = -1.0;
    if (n > 0)
        ck(mplbhip_iterate(S().ctx, criteria, (int)period, (long long)n, &done, res), "Iterate");
    return done;
}

}  // namespace

// An App that never steps (Tests/ConservationTest3D/conservation3d.cpp)
// defines no body-force hook (flowfield.h:114-115).  The reference links it
// because its library is an archive whose evolution object such an App
// never pulls in; this weak definition lets it link here, and an App that
// does step without defining the hook stops with a message.
__attribute__((weak)) void UpdateMacroscopicBodyForce(const Real time) {
    die("Error! the application defines no UpdateMacroscopicBodyForce(Real), which stepping "
        "calls at every step (time %g)", (double)time);
}

// ---------------------------------------------------------------------------
// Block (block.cpp:14-148)
Block::Block(const int blockId, const std::string& blockName,
             const std::vector<int>& blockSize)
    : name(blockName), id(blockId), size(blockSize) {
    if ((int)blockSize.size() != spaceDim)
        die("Error! The SPACEDIM here is inconsistent with the block size");
    block = ops_decl_block(spaceDim, blockName.c_str());
    wholeRange.assign(2 * spaceDim, 0);
    bulkRange.assign(2 * spaceDim, 0);
    for (int a = 0; a < spaceDim; a++) {
        wholeRange[2 * a] = 0;
        wholeRange[2 * a + 1] = size[a];
        bulkRange[2 * a] = 1;
        bulkRange[2 * a + 1] = size[a] - 1;
    }
    for (const auto surface : AllBoundarySurface) {
        std::vector<int> r;
        for (int a = 0; a < spaceDim; a++) {
            r.push_back(RangeStart(a, surface));
            r.push_back(RangeEnd(a, surface));
        }
        boundarySurfaceRange[surface] = r;
    }
}

int Block::RangeStart(const int axis, const BoundarySurface s) const {
    int start = 0;
    if (axis == 0 && (s == BoundarySurface::Right || s == BoundarySurface::RightTop ||
                      s == BoundarySurface::RightBottom
#ifdef OPS_3D
                      || s == BoundarySurface::RightFront || s == BoundarySurface::RightBack
#endif
                      ))
        start = size[axis] - 1;
    // the reference tests the misspelt macro OP_3D here, so TopBack and
    // TopFront keep the full y range
    if (axis == 1 && (s == BoundarySurface::Top || s == BoundarySurface::LeftTop ||
                      s == BoundarySurface::RightTop))
        start = size[axis] - 1;
#ifdef OPS_3D
    if (axis == 2 && (s == BoundarySurface::Front || s == BoundarySurface::LeftFront ||
                      s == BoundarySurface::RightFront || s == BoundarySurface::TopFront ||
                      s == BoundarySurface::BottomFront))
        start = size[axis] - 1;
#endif
    return start;
}

int Block::RangeEnd(const int axis, const BoundarySurface s) const {
    int end = size[axis];
    if (axis == 0 && (s == BoundarySurface::Left || s == BoundarySurface::LeftBottom ||
                      s == BoundarySurface::LeftTop
#ifdef OPS_3D
                      || s == BoundarySurface::LeftFront || s == BoundarySurface::LeftBack
#endif
                      ))
        end = 1;
    if (axis == 1 && (s == BoundarySurface::Bottom || s == BoundarySurface::LeftBottom ||
                      s == BoundarySurface::RightBottom
#ifdef OPS_3D
                      || s == BoundarySurface::BottomBack || s == BoundarySurface::BottomFront
#endif
                      ))
        end = 1;
#ifdef OPS_3D
    if (axis == 2 && (s == BoundarySurface::Back || s == BoundarySurface::LeftBack ||
                      s == BoundarySurface::RightBack || s == BoundarySurface::TopBack ||
                      s == BoundarySurface::BottomBack))
        end = 1;
#endif
    return end;
}

void Block::AddNeighbor(BoundarySurface surface, const Neighbor& neighbor) {
    if (neighbors.find(surface) != neighbors.end())
        die("Error! There is already a neighbor defined for Block %s!", name.c_str());
    neighbors.emplace(surface, neighbor);
}

// ---------------------------------------------------------------------------
// flowfield.h
const BlockGroup& g_Block() { return S().blocks; }
RealField& g_f() { return S().f; }
RealField& g_fStage() { return S().fStage; }
RealFieldGroup& g_MacroVars() { return S().macroVars; }
RealFieldGroup& g_MacroVarsCopy() { return S().macroVarsCopy; }
RealFieldGroup& g_MacroBodyforce() { return S().macroBodyforce; }
RealField& g_CoordinateXYZ() { return S().coordinateXYZ; }
IntFieldGroup& g_NodeType() { return S().nodeType; }
IntField& g_GeometryProperty() { return S().geometryProperty; }
Real TimeStep() { return S().dt; }
const Real* pTimeStep() { return &S().dt; }
const std::string& CaseName() { return S().caseName; }
Real TotalMeshSize() { return 0; }
std::map<int, Real>& g_ResidualError() { return S().residualError; }
bool IsTransient() { return S().transient; }

void SetTimeStep(Real dt) {
    S().dt = dt;
    if (S().ctx) ck(mplbhip_set_timestep(S().ctx, dt), "SetTimeStep");
}

void DefineCase(const std::string& caseName, const int spaceDim, const bool transient) {
#ifdef OPS_3D
    const int dim = 3;
#else
    const int dim = 2;
#endif
    if (spaceDim != dim) die("Error! The SPACEDIM here is inconsistent with the build");
    S().caseName = caseName;
    S().spaceDim = spaceDim;
    S().transient = transient;
}

void DefineBlocks(const std::vector<int>& blockIds, const std::vector<std::string>& blockNames,
                  const std::vector<int>& blockSizes, const Real meshSize,
                  const std::map<int, std::vector<Real>>& startPos) {
    const int D = S().spaceDim;
    if (blockIds.size() != blockNames.size() || blockSizes.size() != D * blockIds.size())
        die("Error! inconsistent block ids, names and sizes");
    if (startPos.size() != blockIds.size())
        die("Error! We expect starting points for %zu blocks, but only received %zu",
            blockIds.size(), startPos.size());
    for (size_t i = 0; i < blockIds.size(); i++) {
        std::vector<int> size(blockSizes.begin() + D * i, blockSizes.begin() + D * (i + 1));
        S().blocks.emplace(blockIds[i], Block(blockIds[i], blockNames[i], size));
    }
    S().meshSize = meshSize;
    for (const auto& idPos : startPos) {
        std::vector<std::vector<Real>> c(D);
        const Block& b = S().blocks.at(idPos.first);
        for (int a = 0; a < D; a++) {
            c[a].resize(b.Size()[a]);
            for (int n = 0; n < b.Size()[a]; n++) c[a][n] = idPos.second.at(a) + n * meshSize;
        }
        S().coordinates[idPos.first] = c;
    }
    S().coordinateXYZ.SetDataDim(D);
    S().coordinateXYZ.CreateFieldFromScratch(S().blocks);
    S().geometryProperty.CreateFieldFromScratch(S().blocks);
}

void DefineBlockConnection(const std::vector<int>& fromBlock,
                           const std::vector<BoundarySurface>& fromSurface,
                           const std::vector<int>& toBlock,
                           const std::vector<BoundarySurface>& toSurface,
                           const std::vector<VertexType>& connectionType) {
    const size_t n = fromBlock.size();
    if (fromSurface.size() != n || toBlock.size() != n || toSurface.size() != n ||
        connectionType.size() != n)
        die("Please input consistent halo pairs!");
    for (size_t i = 0; i < n; i++) {
        Neighbor nb;
        nb.blockId = toBlock[i];
        nb.surface = toSurface[i];
        nb.type = connectionType[i];
        S().blocks.at(fromBlock[i]).AddNeighbor(fromSurface[i], nb);
        S().connections.push_back({fromBlock[i], (int)fromSurface[i], toBlock[i],
                                   (int)toSurface[i], (int)connectionType[i]});
    }
}

void TransferHalos() {}  // inside the device step

void Partition() { PrepareFlowField(); }

static void upload_temperature();

void PrepareFlowField() {
    State& s = S();
    if (!s.ctx) die("Error! please call DefineComponents before Partition");
    const Component& c = the_component();
    for (const auto& l : s.connections)
        ck(mplbhip_define_connection(s.ctx, l[0], l[1], l[2], l[3], l[4]), "connection");
    s.forceMacros = false;
    for (const auto& mv : c.macroVars)
        if (mv.first == Variable_U_Force || mv.first == Variable_V_Force ||
            mv.first == Variable_W_Force)
            s.forceMacros = true;
    // the swap scheme runs the _Swap collision and force kernels
    // (model_wrapper.cpp:49-60, 305-336); either scheme with the other's
    // kernels would mix two lattice layouts, so that is an error here
    const bool swap = s.scheme == Scheme_StreamCollision_Swap;
    if (swap) {
        if (c.collisionType != Collision_BGKIsothermal2nd_Swap)
            die("Error! the swap scheme needs Collision_BGKIsothermal2nd_Swap");
        if (c.bodyForceType != BodyForce_None_Swap && c.bodyForceType != BodyForce_1st_Swap)
            die("Error! the swap scheme needs BodyForce_None_Swap or BodyForce_1st_Swap");
        ck(mplbhip_define_scheme(s.ctx, (int)Scheme_StreamCollision_Swap), "DefineScheme");
    } else {
        if (c.collisionType != Collision_BGKIsothermal2nd &&
            c.collisionType != Collision_BGKThermal4th)
            die("Error! only Collision_BGKIsothermal2nd and Collision_BGKThermal4th are "
                "implemented for Scheme_StreamCollision");
        if (c.collisionType == Collision_BGKThermal4th) {
            if (!c.macroVars.count(Variable_T))
                die("Error! Collision_BGKThermal4th reads the macro variable Variable_T");
            ck(mplbhip_define_collision_type(s.ctx, 1), "DefineCollision (thermal)");
        }
        if (c.bodyForceType != BodyForce_None && c.bodyForceType != BodyForce_1st)
            die("Error! only BodyForce_None and BodyForce_1st are implemented for "
                "Scheme_StreamCollision");
    }
    const int forceType = (c.bodyForceType == BodyForce_1st ||
                           c.bodyForceType == BodyForce_1st_Swap) ? 1 : 0;
    ck(mplbhip_define_collision(s.ctx, c.tauRef, forceType, s.forceMacros), "collision");
    for (const BlockBoundary& bc : s.boundaries) {
        if (bc.componentID != c.id) die("Error! boundary names an undefined component");
        ck(mplbhip_define_boundary(s.ctx, bc.blockIndex, (int)bc.boundarySurface,
                                   (int)bc.boundaryScheme, bc.givenVars.data(),
                                   (int)bc.givenVars.size(), (int)bc.boundaryType),
           "DefineBlockBoundary");
    }
    ck(mplbhip_prepare(s.ctx), "Partition");
    // host mirrors of node type, geometry and coordinates
    for (const auto& idBlock : s.blocks) {
        const int b = idBlock.first;
        const std::vector<int>& n = idBlock.second.Size();
        const int nz = s.spaceDim == 3 ? n[2] : 1;
        std::vector<int> nt((size_t)n[0] * n[1] * nz), vg(nt.size());
        ck(mplbhip_node_info(s.ctx, b, 0, nt.data()), "node_info");
        ck(mplbhip_node_info(s.ctx, b, 1, vg.data()), "node_info");
        size_t p = 0;
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < n[1]; j++)
                for (int i = 0; i < n[0]; i++, p++) {
                    s.nodeType.begin()->second.Value(b, i, j, k) = nt[p];
                    s.geometryProperty.Value(b, i, j, k) = vg[p];
                    for (int a = 0; a < s.spaceDim; a++) {
                        const int ix = a == 0 ? i : (a == 1 ? j : k);
                        s.coordinateXYZ.Value(b, i, j, k, a) = s.coordinates.at(b)[a][ix];
                    }
                }
    }
    s.partitioned = true;
    if (s.restartStep != 0) {
        // the restart state read from the files: macros (with their
        // residual copies) and f, which the first step collides
        for (const auto& idBlock : s.blocks) {
            const int b = idBlock.first;
            std::vector<double> m[4];
            for (int v = 0; v < 4; v++) {
                const int id = v == 0 ? c.macroVars.at(Variable_Rho).id
                                      : (v == 1 ? c.uId : (v == 2 ? c.vId : -1));
#ifdef OPS_3D
                const int idw = v == 3 ? c.wId : id;
#else
                const int idw = id;
#endif
                if (idw >= 0 && s.macroVars.count(idw))
                    m[v] = interior(s.macroVars.at(idw).at(b));
                else
                    m[v].assign(m[0].size(), 0.0);
            }
            ck(mplbhip_set_macros(s.ctx, b, m[0].data(), m[1].data(), m[2].data(),
                                  s.spaceDim == 3 ? m[3].data() : nullptr),
               "restart macros");
            ops_dat d = s.f.at(b);
            const double* p = reinterpret_cast<const double*>(d->data.data());
            const size_t nn = (size_t)d->size[0] * d->size[1] * d->size[2];
            std::vector<double> f((size_t)NUMXI * nn);
            size_t n = 0;
            for (int k = 0; k < d->size[2]; k++)
                for (int j = 0; j < d->size[1]; j++)
                    for (int i = 0; i < d->size[0]; i++, n++)
                        for (int q = 0; q < NUMXI; q++) f[q * nn + n] = p[d->offset(i, j, k) + q];
            ck(mplbhip_set_f(s.ctx, b, f.data()), "restart f");
        }
        upload_temperature();
    }
}

Real GetMaximumResidual(const SizeType checkPeriod) {
    Real maxResError{0};
    for (const auto& error : S().residualError) {
        const Real rel = error.second / (checkPeriod * TimeStep());
        if (maxResError <= rel) maxResError = rel;
    }
    return maxResError;
}

void CalcResidualError() {
    const Component& c = the_component();
    double raw[4] = {0, 0, 0, 0}, mx = 0;
    ck(mplbhip_residual(S().ctx, 1, raw, &mx), "CalcResidualError");
    for (const auto& mv : c.macroVars) {
        const int slot = var_slot(c, mv.second.id);
        if (slot >= 0) S().residualError[mv.second.id] = raw[slot];
    }
    download_macros();
}

void DispResidualError(const int iter, const SizeType checkPeriod) {
    ops_printf("##########Residual Error at %i time step##########\n", iter);
    for (auto& compo : S().components)
        for (auto& mv : compo.second.macroVars) {
            const Real r = S().residualError[mv.second.id] / (checkPeriod * TimeStep());
            ops_printf("Residual of %s = %.17g\n", mv.second.name.c_str(), r);
        }
}

void CopyCurrentMacroVar() {}  // the device keeps MacroVarsCopy itself

// Src/flowfield.cpp:118-137: macros, coordinates and body force; f; node
// type and geometry, each block into <case>_<block>_T<step>.h5
void WriteFlowfieldToHdf5(const SizeType timeStep) {
    if (!hdf5_enabled()) return;
    for (auto& idField : S().macroVars) write_field(idField.second, timeStep);
    write_field(S().coordinateXYZ, timeStep);
    for (auto& idField : S().macroBodyforce) write_field(idField.second, timeStep);
}

void WriteDistributionsToHdf5(const SizeType timeStep) {
    if (!hdf5_enabled()) return;
    State& s = S();
    ensure_f_mirror();
    for (const auto& idBlock : s.blocks) {
        ops_dat d = s.f.at(idBlock.first);
        const size_t nn = (size_t)d->size[0] * d->size[1] * d->size[2];
        std::vector<double> f((size_t)NUMXI * nn);
        ck(mplbhip_get_f(s.ctx, idBlock.first, f.data()), "WriteDistributionsToHdf5");
        double* p = reinterpret_cast<double*>(d->data.data());
        size_t n = 0;
        for (int k = 0; k < d->size[2]; k++)
            for (int j = 0; j < d->size[1]; j++)
                for (int i = 0; i < d->size[0]; i++, n++)
                    for (int q = 0; q < NUMXI; q++) p[d->offset(i, j, k) + q] = f[q * nn + n];
    }
    write_field(s.f, timeStep);
}

void WriteNodePropertyToHdf5(const SizeType timeStep) {
    if (!hdf5_enabled()) return;
    write_field(S().geometryProperty, timeStep);
    for (auto& idField : S().nodeType) write_field(idField.second, timeStep);
}

// ---------------------------------------------------------------------------
// model.h
const std::map<int, Component>& g_Components() { return S().components; }

void DefineComponents(const std::vector<std::string>& compoNames,
                      const std::vector<int>& compoId,
                      const std::vector<std::string>& lattNames,
                      const std::vector<Real> tauRef, const SizeType timeStep) {
    State& s = S();
    if (s.blocks.empty()) die("Error:please call DefineBlock first!");
    if (compoNames.size() != 1)
        die("Error! this build runs one component per case (got %zu)", compoNames.size());
    const std::string latt = lattNames.at(0);
    if ((s.spaceDim == 2 && latt != "d2q9") ||
        (s.spaceDim == 3 && latt != "d3q19" && latt != "d3q15"))
        die("Error! lattice %s is not implemented (d2q9 in 2D, d3q19 or d3q15 in 3D)",
            latt.c_str());
    const bool q15 = latt == "d3q15";
    Component comp;
    comp.id = compoId.at(0);
    comp.name = compoNames.at(0);
    comp.latticeName = latt;
    comp.tauRef = tauRef.at(0);
    NUMCOMPONENTS = 1;
    NUMXI = s.spaceDim == 2 ? 9 : (q15 ? 15 : 19);
    LATTDIM = s.spaceDim;
    CS = std::sqrt(3.0);
    comp.index[0] = 0;
    comp.index[1] = NUMXI - 1;
    s.components.emplace(comp.id, comp);
    // host copies of the tables (model.cpp:99-150)
    static const int c2[9][2] = {{0, 0},  {-1, 1}, {-1, 0}, {-1, -1}, {0, -1},
                                 {1, -1}, {1, 0},  {1, 1},  {0, 1}};
    static const int c15[15][3] = {
        {0, 0, 0},   {-1, 0, 0},  {0, -1, 0}, {0, 0, -1}, {-1, -1, -1},
        {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1}, {1, 0, 0},  {0, 1, 0},
        {0, 0, 1},   {1, 1, 1},   {1, 1, -1}, {1, -1, 1}, {1, -1, -1}};
    static const int c3[19][3] = {
        {0, 0, 0},  {-1, 0, 0},  {0, -1, 0}, {0, 0, -1}, {-1, -1, 0},
        {-1, 1, 0}, {-1, 0, -1}, {-1, 0, 1}, {0, -1, -1}, {0, -1, 1},
        {1, 0, 0},  {0, 1, 0},   {0, 0, 1},  {1, 1, 0},  {1, -1, 0},
        {1, 0, 1},  {1, 0, -1},  {0, 1, 1},  {0, 1, -1}};
    delete[] XI;
    delete[] WEIGHTS;
    delete[] OPP;
    XI = new Real[NUMXI * LATTDIM];
    WEIGHTS = new Real[NUMXI];
    OPP = new int[NUMXI];
    for (int q = 0; q < NUMXI; q++) {
        for (int a = 0; a < LATTDIM; a++)
            XI[q * LATTDIM + a] = s.spaceDim == 2 ? c2[q][a] : (q15 ? c15[q][a] : c3[q][a]);
        if (s.spaceDim == 2) {
            WEIGHTS[q] = q == 0 ? 4.0 / 9.0 : (q % 2 == 1 ? 1.0 / 36.0 : 1.0 / 9.0);
        } else if (q15) {
            WEIGHTS[q] = q == 0 ? 2.0 / 9.0 : ((q <= 3 || (q >= 8 && q <= 10)) ? 1.0 / 9.0 : 1.0 / 72.0);
        } else {
            WEIGHTS[q] = q == 0 ? 1.0 / 3.0 : ((q <= 3 || (q >= 10 && q <= 12)) ? 1.0 / 18.0 : 1.0 / 36.0);
        }
    }
    for (int q = 0; q < NUMXI; q++)
        for (int r = 0; r < NUMXI; r++) {
            bool rev = true;
            for (int a = 0; a < LATTDIM; a++) rev = rev && XI[q * LATTDIM + a] + XI[r * LATTDIM + a] == 0;
            if (rev) { OPP[q] = r; break; }
        }
    XIMAXVALUE = CS;
    const char* dev = std::getenv("MPLB_DEVICE");
    const int device = dev ? std::atoi(dev) : 0;
    ck(mplbhip_create(&s.ctx, device, s.spaceDim, storage_from_env()), "mplbhip_create");
    if (!s.ctx) die("Error! no GPU context");
    if (q15) ck(mplbhip_define_lattice(s.ctx, 15), "DefineComponents (d3q15)");
    for (const auto& idBlock : s.blocks) {
        int size[3] = {1, 1, 1};
        for (int a = 0; a < s.spaceDim; a++) size[a] = idBlock.second.Size()[a];
        ck(mplbhip_define_block(s.ctx, idBlock.first, size), "define_block");
    }
    IntField nt{"NodeType_" + comp.name};
    nt.CreateFieldFromScratch(s.blocks);
    s.nodeType.emplace(comp.id, nt);
    // a restart reads f of every block from <case>_<block>_T<step>.h5
    // (model.cpp:299-311); it reaches the device at Partition
    s.restartStep = timeStep;
    if (timeStep != 0) {
        ensure_f_mirror();
        read_field(s.f, timeStep);
    }
}

void DefineMacroVars(std::vector<VariableTypes> types, std::vector<std::string> names,
                     std::vector<int> varId, std::vector<int> compoId, const SizeType timeStep) {
    State& s = S();
    if (s.components.empty()) die("Error:please call DefineComponent first!");
    for (size_t i = 0; i < names.size(); i++) {
        MacroVariable mv;
        mv.name = names.at(i);
        mv.id = varId.at(i);
        mv.type = types.at(i);
        s.components.at(compoId.at(i)).macroVars.emplace(mv.type, mv);
        RealField field{mv.name};
        field.CreateFieldFromScratch(s.blocks);
        // a restart reads the macros from the files (model.cpp:379-388)
        if (timeStep != 0) read_field(field, timeStep);
        s.macroVars.emplace(mv.id, field);
        if (!s.transient) {
            RealField copy{mv.name + "Copy"};
            copy.CreateFieldFromScratch(s.blocks);
            s.macroVarsCopy.emplace(mv.id, copy);
            s.residualError.emplace(mv.id, 0.0);
        }
    }
    for (auto& idCompo : s.components) {
        Component& c = idCompo.second;
        const auto& m = c.macroVars;
        auto pick = [&](VariableTypes a, VariableTypes b) {
            if (m.count(b)) return m.at(b).id;
            if (m.count(a)) return m.at(a).id;
            return 0;
        };
        c.uId = pick(Variable_U, Variable_U_Force);
        c.vId = pick(Variable_V, Variable_V_Force);
#ifdef OPS_3D
        c.wId = pick(Variable_W, Variable_W_Force);
#endif
    }
}

void DefineCollision(std::vector<CollisionType> types, std::vector<int> compoId) {
    if (types.size() != compoId.size()) die("Error! collision types and components differ");
    for (size_t i = 0; i < types.size(); i++)
        S().components.at(compoId[i]).collisionType = types[i];
}

void DefineBodyForce(std::vector<BodyForceType> types, std::vector<SizeType> compoId) {
    if (types.size() != compoId.size()) die("Error! force types and components differ");
    for (size_t i = 0; i < types.size(); i++) {
        Component& c = S().components.at((int)compoId[i]);
        c.bodyForceType = types[i];
        RealField force{"Force_" + c.name};
        force.SetDataDim(S().spaceDim);
        force.CreateFieldFromScratch(S().blocks);
        S().macroBodyforce.emplace(c.id, force);
    }
}

void DefineInitialCondition(std::vector<InitialType> types, std::vector<int> compoId) {
    for (size_t i = 0; i < types.size(); i++)
        S().components.at(compoId[i]).initialType = types[i];
}

// the Variable_T field of every block, which the thermal collision reads
// (KerCollideBGKThermal[3D] takes it as it is: nothing updates it)
static void upload_temperature() {
    State& s = S();
    const Component& c = the_component();
    if (c.collisionType != Collision_BGKThermal4th) return;
    const int id = c.macroVars.at(Variable_T).id;
    for (const auto& idBlock : s.blocks) {
        const int b = idBlock.first;
        const std::vector<double> t = interior(s.macroVars.at(id).at(b));
        ck(mplbhip_set_temperature(s.ctx, b, t.data()), "Variable_T upload");
    }
}

static void initial_condition() {
    State& s = S();
    const Component& c = the_component();
    if (c.initialType != Initial_BGKFeq2nd) die("Error! unknown initial condition");
    const int ids[4] = {c.macroVars.at(Variable_Rho).id, c.uId, c.vId,
#ifdef OPS_3D
                        c.wId
#else
                        -1
#endif
    };
    for (const auto& idBlock : s.blocks) {
        const int b = idBlock.first;
        std::vector<double> v[4];
        for (int k = 0; k < 1 + s.spaceDim; k++) v[k] = interior(s.macroVars.at(ids[k]).at(b));
        ck(mplbhip_set_macros(s.ctx, b, v[0].data(), v[1].data(), v[2].data(),
                              s.spaceDim == 3 ? v[3].data() : nullptr),
           "SetInitialMacrosVars upload");
        for (int k = 0; k < 1 + s.spaceDim; k++) s.macroVars.at(ids[k]).at(b)->dirty = false;
    }
    upload_force_if_dirty();
    upload_temperature();
    ck(mplbhip_init_feq(s.ctx), "PreDefinedInitialCondition");
}

#ifdef OPS_3D
void UpdateMacroVars3D() {
    ck(mplbhip_update_macros(S().ctx), "UpdateMacroVars3D");
    download_macros();
}
void PreDefinedInitialCondition3D() { initial_condition(); }
void PreDefinedBodyForce3D() {}
void PreDefinedCollision3D() {}
void PredefinedStream3D() {}
void ImplementBoundary3D() {}
#endif
#ifdef OPS_2D
void UpdateMacroVars() {
    ck(mplbhip_update_macros(S().ctx), "UpdateMacroVars");
    download_macros();
}
void PreDefinedInitialCondition() { initial_condition(); }
void PreDefinedBodyForce() {}
void PreDefinedCollision() {}
void Stream() {}
void ImplementBoundary() {}
#endif

void DestroyModel() {
    delete[] XI;
    delete[] WEIGHTS;
    delete[] OPP;
    XI = nullptr;
    WEIGHTS = nullptr;
    OPP = nullptr;
}

// ---------------------------------------------------------------------------
// scheme.h
static ops_stencil make_stencil(const char* name) { return ops_decl_stencil(0, 0, nullptr, name); }
ops_stencil LOCALSTENCIL{make_stencil("local")};
ops_stencil ONEPTREGULARSTENCIL{make_stencil("onept")};
ops_stencil TWOPTREGULARSTENCIL{make_stencil("twopt")};
ops_stencil ONEPTLATTICESTENCIL{make_stencil("lattice")};
static int schemeHaloPt{1};
void SetupCommonStencils() {}
SchemeType Scheme() { return S().scheme; }
int SchemeHaloNum() { return schemeHaloPt; }
void SetSchemeHaloNum(const int n) { schemeHaloPt = n; }
void DefineScheme(const SchemeType scheme) {
    if (scheme != Scheme_StreamCollision && scheme != Scheme_StreamCollision_Swap)
        die("Error! only Scheme_StreamCollision and Scheme_StreamCollision_Swap are "
            "implemented in this build");
    S().scheme = scheme;
    if (scheme == Scheme_StreamCollision_Swap)
        ops_printf("The swap stream-collision scheme is chosen!\n");
    else
        ops_printf("The stream-collision scheme is chosen!\n");
}

// ---------------------------------------------------------------------------
// boundary.h
static int boundaryHaloPt{1};
int BoundaryHaloNum() { return boundaryHaloPt; }
void SetBoundaryHaloNum(const int n) { boundaryHaloPt = n; }
const std::vector<BlockBoundary>& BlockBoundaries() { return S().boundaries; }

void DefineBlockBoundary(int blockIndex, int componentID, BoundarySurface boundarySurface,
                         BoundaryScheme boundaryScheme,
                         const std::vector<VariableTypes>& macroVarTypes,
                         const std::vector<Real>& macroVarValues,
                         const VertexType boundaryType) {
    if (macroVarTypes.size() != macroVarValues.size())
        ops_printf("Error: Please specify same number of variable types and values\n");
    BlockBoundary bc;
    bc.blockIndex = blockIndex;
    bc.componentID = componentID;
    bc.givenVars = macroVarValues;
    bc.boundarySurface = boundarySurface;
    bc.boundaryScheme = boundaryScheme;
    bc.macroVarTypesatBoundary = macroVarTypes;
    bc.boundaryType = boundaryType;
    S().boundaries.push_back(bc);
    ops_printf("The scheme %i is adopted for Component %i at Surface %i, boundary type %i "
               "and Block %i\n",
               (int)boundaryScheme, componentID, (int)boundarySurface, (int)boundaryType,
               blockIndex);
}

void DefineBlockBoundary(int blockIndex, int componentID, BoundarySurface boundarySurface,
                         const VertexType boundaryType) {
    if (boundaryType != VertexType::VirtualBoundary)
        ops_printf("Error: This routine is specially for defining virtual boundary!\n");
    DefineBlockBoundary(blockIndex, componentID, boundarySurface, BoundaryScheme::None, {}, {},
                        boundaryType);
}

// ---------------------------------------------------------------------------
// evolution.h
void StreamCollision(const Real time) {
    UpdateMacroscopicBodyForce(time);
    upload_force_if_dirty();
    ck(mplbhip_step(S().ctx, 1), "StreamCollision");
}

// SwapStreamCollision (evolution.cpp:195-252): macros from f, collision and
// force in place, halo exchange of f, local swap, swap stream, boundaries --
// one mplbhip_step of a context defined with the swap scheme.
void SwapStreamCollision(const Real time) {
    if (S().scheme != Scheme_StreamCollision_Swap)
        die("Error! SwapStreamCollision needs DefineScheme(Scheme_StreamCollision_Swap)");
    UpdateMacroscopicBodyForce(time);
    upload_force_if_dirty();
    ck(mplbhip_step(S().ctx, 1), "SwapStreamCollision");
}

namespace hilemms_detail {
// Either cycle function hands the loop to the library, which runs the scheme
// the context was defined with; the cycle must match that scheme.
bool IsStreamCollision(void (*cycle)(Real)) {
    const bool swap = S().scheme == Scheme_StreamCollision_Swap;
    if (cycle == &StreamCollision && swap)
        die("Error! StreamCollision given to a case defined with the swap scheme");
    if (cycle == &SwapStreamCollision && !swap)
        die("Error! SwapStreamCollision given to a stream-collision case");
    return cycle == &StreamCollision || cycle == &SwapStreamCollision;
}
void CheckpointUpdate(SizeType iter) {
    ops_printf("%d iterations!\n", (int)iter);
#ifdef OPS_3D
    UpdateMacroVars3D();
#else
    UpdateMacroVars();
#endif
    WriteFlowfieldToHdf5(iter);
    WriteDistributionsToHdf5(iter);
    WriteNodePropertyToHdf5(iter);
}
Real CheckpointResidual(SizeType iter, SizeType checkPeriod) {
#ifdef OPS_3D
    UpdateMacroVars3D();
#else
    UpdateMacroVars();
#endif
    CalcResidualError();
    const Real r = GetMaximumResidual(checkPeriod);
    DispResidualError((int)iter, checkPeriod);
    WriteFlowfieldToHdf5(iter);
    WriteDistributionsToHdf5(iter);
    WriteNodePropertyToHdf5(iter);
    return r;
}
void Finish() {
    ops_printf("Simulation finished! Exiting...\n");
    DestroyModel();
}
}  // namespace hilemms_detail

// Iterate(steps, checkPointPeriod) (evolution.cpp:71-100): runs on the
// device between checkpoints (with the force hook at every step's time);
// macros come back at every checkpoint.
void Iterate(const SizeType steps, const SizeType checkPointPeriod, const SizeType start) {
    ops_printf("Starting the iteration...\n");
    upload_force_if_dirty();
    SizeType iter = start;
    while (iter < start + steps) {
        const SizeType to_check = checkPointPeriod - (iter % checkPointPeriod);
        const SizeType n = std::min<SizeType>(to_check, start + steps - iter);
        double res = 0;
        advance(iter, n, -1.0, checkPointPeriod, &res, iter == start);
        iter += n;
        if (iter % checkPointPeriod == 0) {
            ops_printf("%d iterations!\n", (int)iter);
            download_macros();
            WriteFlowfieldToHdf5(iter);
            WriteDistributionsToHdf5(iter);
            WriteNodePropertyToHdf5(iter);
        }
    }
    hilemms_detail::Finish();
}

// Iterate(convergenceCriteria, checkPointPeriod) (evolution.cpp:102-135)
void Iterate(const Real convergenceCriteria, const SizeType checkPointPeriod,
             const SizeType start) {
    ops_printf("Starting the iteration...\n");
    State& s = S();
    upload_force_if_dirty();
    SizeType iter = start;
    Real residualError{1};
    do {
        double res = -1;
        // up to the next check step (the loop's first pass may start
        // between two of them)
        const SizeType n = checkPointPeriod - (iter % checkPointPeriod);
        iter += (SizeType)advance(iter, n, convergenceCriteria, checkPointPeriod, &res,
                                  iter == start);
        if (iter % checkPointPeriod == 0 && res >= -0.5) {
            const Component& c = the_component();
            double raw[4] = {0, 0, 0, 0}, mx = 0;
            // the residual ratios of the check step just taken
            ck(mplbhip_last_residual(s.ctx, raw, &mx), "residual");
            for (const auto& mv : c.macroVars) {
                const int slot = var_slot(c, mv.second.id);
                if (slot >= 0) s.residualError[mv.second.id] = raw[slot];
            }
            residualError = GetMaximumResidual(checkPointPeriod);
            download_macros();
            DispResidualError((int)iter, checkPointPeriod);
            WriteFlowfieldToHdf5(iter);
            WriteDistributionsToHdf5(iter);
            WriteNodePropertyToHdf5(iter);
        }
    } while (residualError >= convergenceCriteria);
    hilemms_detail::Finish();
}
